// Reader of hfscf's .dat input files (see input.hpp).
#include "input.hpp"

#include <algorithm>
#include <cctype>
#include <fstream>
#include <sstream>

namespace hfscf {

namespace {

// Keyword types: a list of accepted values ("true"/"false" for the boolean ones), or one of the
// value kinds below. The names, kinds and accepted values are hfscf's keyword table
// (hfscf: src/molecule/hfscf_keywords.hpp:14-72).
constexpr const char* kAny = "*";
constexpr const char* kInt = "int";
constexpr const char* kDouble = "double";

const std::multimap<std::string, std::string>& keyword_table() {
    static const std::multimap<std::string, std::string> t = [] {
        std::multimap<std::string, std::string> m;
        auto one_of = [&](const char* key, std::initializer_list<const char*> vals) {
            for (const char* v : vals) m.emplace(key, v);
        };
        auto boolean = [&](const char* key) { one_of(key, {"true", "false"}); };
        for (const char* k : {"basis_set", "point_group_override"}) m.emplace(k, kAny);
        one_of("post_scf", {"MP2", "MP3", "CCSD", "CCSD(T)"});
        one_of("scf_type", {"RHF", "UHF"});
        one_of("initial_guess", {"CORE", "SAD"});
        one_of("gradient_type", {"SCF", "MP2"});
        one_of("ci_type", {"CIS", "RPA"});
        one_of("geom_opt_tol", {"VERYHIGH", "HIGH", "MEDIUM", "LOW"});
        one_of("geom_opt", {"SCF", "MP2"});
        one_of("geom_opt_algorithm", {"RFO", "CGSD"});
        one_of("frequencies", {"SCF", "MP2"});
        one_of("units", {"bohr", "angstrom"});
        one_of("geom_opt_guess_hessian", {"SCHLEGEL", "SIMPLE"});
        one_of("point_group_threshold", {"TIGHT", "DEFAULT", "RELAXED"});
        for (const char* k : {"electronic_properties", "mol_props_cphf_iter", "use_puream", "align_geom", "use_symmetry",
                              "geom_opt_constrain_angles", "geom_opt_constrain_dihedrals", "geom_opt_constrain_oop",
                              "soscf", "hess_project_translations_only", "geom_opt_write_xyz", "freq_write_molden",
                              "eri_screen", "freeze_core", "uhf_guess_mix", "scf_direct", "symmetrize_geom"})
            boolean(k);
        for (const char* k : {"diis_range", "ccsd_diis_range", "max_ccsd_iter", "max_scf_iter", "max_sad_iter",
                              "max_soscf_iter", "max_geomopt_iter", "charge", "multiplicity"})
            m.emplace(k, kInt);
        for (const char* k : {"ccsd_energy_tol", "energy_tol", "sad_energy_tol", "ccsd_rms_tol", "rms_tol", "sad_rms_tol",
                              "soscf_rms_tol", "density_damp", "integral_tol", "geom_opt_stepsize",
                              "grad_step_from_energy", "hessian_step_from_grad", "hessian_step_from_energy",
                              "thermo_chem_pressure", "thermo_chem_temperature"})
            m.emplace(k, kDouble);
        return m;
    }();
    return t;
}

// Element symbols H..Kr (atomic number = index + 1).
const char* const kSymbols[] = {"H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na", "Mg",
                                "Al", "Si", "P",  "S",  "Cl", "Ar", "K",  "Ca", "Sc", "Ti", "V",  "Cr",
                                "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr"};
constexpr double kAngstromToBohr = 1.8897261254535;  // hfscf: src/molecule/hfscf_constants.hpp:5
constexpr int kMaxAtomLines = 100;                    // hfscf_molecule.cpp:99-108
constexpr int kMaxKeywordLines = 1000;                // hfscf_molecule.cpp:66-72

std::string strip_blanks(const std::string& s) {
    std::string r;
    for (char c : s)
        if (c != ' ' && c != '\t') r += c;
    return r;
}

// The text before the first '=' without blanks (hfscf's get_keyword).
std::string keyword_of(const std::string& line) { return strip_blanks(line.substr(0, line.find('='))); }

bool is_double(const std::string& v) {
    if (v.empty() || v.find_first_not_of("+-0123456789.Ee") != std::string::npos) return false;
    std::istringstream is(v);
    double d;
    is >> d;
    return !is.fail() && is.eof();
}

bool is_int(const std::string& v) {
    size_t k = (!v.empty() && (v[0] == '+' || v[0] == '-')) ? 1 : 0;
    return k < v.size() && v.find_first_not_of("0123456789", k) == std::string::npos;
}

}  // namespace

std::string basis_file_name(const std::string& name) {
    std::string out;
    for (char c : name) {
        if (c == '*')
            out += "_st_";
        else
            out += (char)std::tolower((unsigned char)c);
    }
    return out + ".gbs";
}

bool read_input(const std::string& path, InputFile& out, std::string& err) {
    std::ifstream f(path);
    if (!f) {
        err = "cannot open input file " + path;
        return false;
    }
    out = InputFile{};
    std::string line;
    std::getline(f, line);
    if (line.substr(0, 1) != "#") {
        err = "Invalid comment line, expecting \"# string\": the first line must be a description or comment line";
        return false;
    }
    out.comment = line.substr(1);
    out.comment.erase(0, out.comment.find_first_not_of(' ') == std::string::npos ? out.comment.size()
                                                                                    : out.comment.find_first_not_of(' '));
    // keyword lines up to the line that opens the atom block
    const auto& table = keyword_table();
    for (int count = 0;;) {
        if (!std::getline(f, line)) {
            err = "missing an opening \"{\" close \"}\" section in the input file";
            return false;
        }
        const std::string key = keyword_of(line);
        if (!key.empty() && key[0] == '{') break;
        if (line.empty() || line[0] == '#' || line.find_first_not_of(" \t") == std::string::npos) continue;
        auto range = table.equal_range(key);
        if (range.first == range.second) {
            err = "Unknown keyword. Line: \"" + line + "\"";
            return false;
        }
        const size_t eq = line.find('=');
        std::string val = eq == std::string::npos ? "" : line.substr(eq + 1);
        val = strip_blanks(val.substr(0, val.find('=')));
        const std::string& kind = range.first->second;
        bool ok;
        if (kind == kAny)
            ok = true;
        else if (kind == kInt)
            ok = is_int(val);
        else if (kind == kDouble)
            ok = is_double(val);
        else {
            ok = false;
            for (auto it = range.first; it != range.second; ++it) ok = ok || it->second == val;
        }
        if (!ok) {
            err = "Invalid keyword specification for " + key + ". Line: \"" + line + "\"";
            return false;
        }
        if (key == "charge")
            out.charge = std::stoi(val);
        else if (key == "multiplicity")
            out.multiplicity = std::stoi(val);
        out.settings[key] = val;
        if (++count == kMaxKeywordLines) {
            err = "Parsing input file failed. Last known line: " + line;
            return false;
        }
    }
    auto units = out.settings.find("units");
    const double scale = (units != out.settings.end() && units->second == "angstrom") ? kAngstromToBohr : 1.0;
    auto basis = out.settings.find("basis_set");
    if (basis != out.settings.end()) out.basis_file = basis_file_name(basis->second);
    // the atom block: "Z-or-symbol x y z" per line up to "}"
    bool closed = false;
    for (int count = 0; count <= kMaxAtomLines && std::getline(f, line); ++count) {
        if (line.find_first_not_of(" \t") == std::string::npos) continue;
        if (line.find('}') != std::string::npos) {
            closed = true;
            break;
        }
        std::istringstream data(line);
        std::string sym, x, y, z;
        data >> sym >> x;
        if (x.empty()) {
            err = "z-matrix geometry is not supported by this reader; give Cartesian coordinates. Line: \"" + line + "\"";
            return false;
        }
        data >> y >> z;
        for (const std::string* v : {&x, &y, &z})
            if (!is_double(*v)) {
                err = "Bad coordinate format on line containing: " + line;
                return false;
            }
        if (sym == "X") continue;  // dummy atom
        int Zn = 0;
        if (std::isdigit((unsigned char)sym[0])) {
            Zn = std::atoi(sym.c_str());
        } else {
            for (int k = 0; k < (int)(sizeof(kSymbols) / sizeof(kSymbols[0])); ++k)
                if (sym == kSymbols[k]) Zn = k + 1;
        }
        if (Zn <= 0) {
            err = "Unsupported element. Line: \"" + line + "\"";
            return false;
        }
        out.Z.push_back(Zn);
        out.xyz.push_back(std::stod(x) * scale);
        out.xyz.push_back(std::stod(y) * scale);
        out.xyz.push_back(std::stod(z) * scale);
    }
    if (!closed) {
        err = "Maximum number of supported atoms is 100, or missing an opening \"{\" close \"}\" section in the input file";
        return false;
    }
    if (out.Z.empty()) {
        err = "No atoms found";
        return false;
    }
    return true;
}

}  // namespace hfscf
