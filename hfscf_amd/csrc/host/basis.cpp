#include "basis.hpp"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace hfscf {

namespace {

const char* kSymbols[] = {"X",  "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",
                          "Ne", "Na", "Mg", "Al", "Si", "P",  "S",  "Cl", "Ar"};

double binom(int n, int k) {
    if (k < 0 || k > n) return 0.0;
    double r = 1.0;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}

double factorial(int n) {
    double r = 1.0;
    for (int i = 2; i <= n; ++i) r *= i;
    return r;
}

int double_factorial(int n) {  // (-1)!! = 1
    int r = 1;
    for (int i = n; i > 1; i -= 2) r *= i;
    return r;
}

int shell_type(const std::string& t) {
    static const std::string kTypes = "SPDFGH";
    if (t.size() != 1) return -1;
    auto p = kTypes.find(t[0]);
    return p == std::string::npos ? -1 : (int)p;
}

}  // namespace

int element_z(const std::string& symbol) {
    for (int z = 1; z < (int)(sizeof(kSymbols) / sizeof(kSymbols[0])); ++z)
        if (symbol == kSymbols[z]) return z;
    throw std::runtime_error("unsupported element: " + symbol);
}

BasisLibrary read_gbs(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open basis file: " + path);
    BasisLibrary lib;
    std::string line;
    std::getline(in, line);
    {
        std::string h;
        for (char ch : line)
            if (ch != ' ' && ch != '\r') h.push_back(ch);
        lib.spherical = (h == "spherical");
    }
    int z = -1;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '!') continue;
        if (line[0] == '*') {
            z = -1;
            continue;
        }
        std::istringstream ss(line);
        std::string tok;
        ss >> tok;
        if (z < 0) {  // element header "O     0"
            z = element_z(tok);
            lib.by_z[z];
            continue;
        }
        // shell header "S   9   1.00" / "SP  3   1.00"
        int nprim = 0;
        ss >> nprim;
        std::vector<double> a(nprim), c1(nprim), c2(nprim);
        for (int i = 0; i < nprim; ++i) {
            std::string pl;
            std::getline(in, pl);
            for (auto& ch : pl)
                if (ch == 'D' || ch == 'd') ch = 'E';
            std::istringstream ps(pl);
            ps >> a[i] >> c1[i];
            if (tok == "SP") ps >> c2[i];
        }
        if (tok == "SP") {
            lib.by_z[z].push_back({0, a, c1});
            lib.by_z[z].push_back({1, a, c2});
        } else {
            int L = shell_type(tok);
            if (L < 0) throw std::runtime_error("unsupported shell type " + tok + " in " + path);
            lib.by_z[z].push_back({L, a, c1});
        }
    }
    return lib;
}

Shell make_shell(int L, int atom, const std::array<double, 3>& r, const std::vector<double>& alpha,
                 const std::vector<double>& coef, int idx, int ids) {
    Shell s;
    s.L = L;
    s.atom = atom;
    s.r = r;
    s.alpha = alpha;
    s.c_raw = coef;
    s.idx = idx;
    s.ids = ids;
    const double m = L + 1.5;
    const double m2 = 0.5 * m;
    const int K = (int)alpha.size();
    double sum = 0.0;
    for (int j = 0; j < K; ++j)
        for (int k = 0; k < K; ++k)
            sum += coef[j] * coef[k] * std::pow(alpha[j] * alpha[k], m2) / std::pow(alpha[j] + alpha[k], m);
    const double norm = 1.0 / std::sqrt(sum * std::pow(M_PI, 1.5) * double_factorial(2 * L - 1) / std::pow(2.0, L));
    s.c.resize(K);
    for (int j = 0; j < K; ++j) s.c[j] = coef[j] * norm * std::pow(alpha[j], m2);
    return s;
}

// Real solid harmonics in terms of x^L-normalised Cartesian Gaussians:
// S_lm = N_lm sum_k (-1)^k C(l,k) C(2(l-k),l) [(l-2k)!/(l-2k-|m|)!] r^(2k) z^(l-2k-|m|) {Re,Im}(x+iy)^|m|
// with N_lm = 2^-l, times sqrt(2 (l-|m|)!/(l+|m|)!) for m != 0.
std::vector<double> ylm_matrix(int l) {
    const int nc = ncart(l), ns = nsph(l);
    std::vector<double> T(nc * ns, 0.0);
    auto cart_index = [](int ly, int lz) {
        int i = ly + lz;
        return i * (i + 1) / 2 + lz;
    };
    for (int mval = -l; mval <= l; ++mval) {
        const int m = std::abs(mval);
        double pre = std::pow(2.0, -l);
        if (m != 0) pre *= std::sqrt(2.0 * factorial(l - m) / factorial(l + m));
        for (int k = 0; k <= (l - m) / 2; ++k) {
            double f = ((k % 2) ? -1.0 : 1.0) * binom(l, k) * binom(2 * (l - k), l);
            if (m != 0) f *= factorial(l - 2 * k) / factorial(l - 2 * k - m);
            f *= pre;
            // r^2k = sum_a C(k,a) x^(2(k-a)) sum_b C(a,b) y^(2(a-b)) z^(2b)
            for (int a = 0; a <= k; ++a)
                for (int b = 0; b <= a; ++b) {
                    const double fab = binom(k, a) * binom(a, b) * f;
                    const int yexp = 2 * (a - b);
                    const int zexp = 2 * b + l - 2 * k - m;
                    // (x + i y)^m = sum_p C(m,p) x^p (i y)^(m-p)
                    for (int p = 0; p <= m; ++p) {
                        const int q = m - p;  // power of (i y)
                        double re = 0, im = 0;
                        switch (q % 4) {
                            case 0: re = 1; break;
                            case 1: im = 1; break;
                            case 2: re = -1; break;
                            case 3: im = -1; break;
                        }
                        const double w = (mval > 0) ? re : (mval < 0 ? im : (q == 0 ? 1.0 : 0.0));
                        if (w == 0.0) continue;
                        T[cart_index(yexp + q, zexp) * ns + (l + mval)] += w * binom(m, p) * fab;
                    }
                }
        }
    }
    return T;
}

int Molecule::nelectrons(int charge) const {
    int n = -charge;
    for (int z : Z) n += z;
    return n;
}

double Molecule::nuclear_repulsion() const {
    double e = 0.0;
    for (int a = 0; a < natom(); ++a)
        for (int b = 0; b < a; ++b) {
            double dx = xyz[a][0] - xyz[b][0], dy = xyz[a][1] - xyz[b][1], dz = xyz[a][2] - xyz[b][2];
            e += Z[a] * Z[b] / std::sqrt(dx * dx + dy * dy + dz * dz);
        }
    return e;
}

void Molecule::set_geometry(const std::vector<std::array<double, 3>>& xyz_bohr) {
    if ((int)xyz_bohr.size() != natom()) throw std::runtime_error("set_geometry: atom count mismatch");
    xyz = xyz_bohr;
    for (auto& s : shells) s.r = xyz[s.atom];
}

Molecule build_molecule(const std::vector<int>& Z, const std::vector<std::array<double, 3>>& xyz_bohr,
                        const BasisLibrary& lib, int pure_override) {
    Molecule mol;
    mol.Z = Z;
    mol.xyz = xyz_bohr;
    mol.pure = pure_override < 0 ? lib.spherical : (pure_override != 0);
    int idx = 0, ids = 0;
    for (int a = 0; a < (int)Z.size(); ++a) {
        auto it = lib.by_z.find(Z[a]);
        if (it == lib.by_z.end()) throw std::runtime_error("element Z=" + std::to_string(Z[a]) + " not in basis");
        for (const auto& g : it->second) {
            mol.shells.push_back(make_shell(g.L, a, xyz_bohr[a], g.alpha, g.coef, idx, ids));
            idx += ncart(g.L);
            ids += nsph(g.L);
        }
    }
    mol.ncart_total = idx;
    mol.nsph_total = ids;
    return mol;
}

void read_xyz(const std::string& path, std::vector<int>& Z, std::vector<std::array<double, 3>>& xyz_bohr) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open " + path);
    int n = 0;
    std::string line;
    std::getline(in, line);
    n = std::stoi(line);
    std::getline(in, line);
    Z.clear();
    xyz_bohr.clear();
    for (int i = 0; i < n; ++i) {
        std::getline(in, line);
        std::istringstream ss(line);
        std::string sym;
        double x, y, z;
        ss >> sym >> x >> y >> z;
        Z.push_back(element_z(sym));
        xyz_bohr.push_back({x * kAngstromToBohr, y * kAngstromToBohr, z * kAngstromToBohr});
    }
}

}  // namespace hfscf
