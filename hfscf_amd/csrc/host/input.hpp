// Reader of hfscf's .dat input files: a comment line, "keyword = value" lines checked against
// hfscf's keyword table, and a { } block of atoms in Cartesian form (hfscf:
// src/molecule/hfscf_molecule.cpp:21-200, src/molecule/hfscf_parser.hpp:48-151,
// src/molecule/hfscf_keywords.hpp). Errors are returned with hfscf's messages instead of exiting.
#pragma once
#include <map>
#include <string>
#include <vector>

namespace hfscf {

struct InputFile {
    std::string comment;                          // the first line without '#' and leading spaces
    std::map<std::string, std::string> settings;  // every keyword given, with its value
    std::vector<int> Z;                           // atoms (dummy atoms X dropped)
    std::vector<double> xyz;                      // natom x 3, bohr
    int charge = 0;
    int multiplicity = 1;
    // basis file name from basis_set: lowercase, '*' -> "_st_", ".gbs" appended (empty if not given)
    std::string basis_file;
};

// Parses `path`; false with the message in err on an error.
bool read_input(const std::string& path, InputFile& out, std::string& err);

// hfscf's basis file name of a basis-set name (hfscf_parser.hpp:134-151).
std::string basis_file_name(const std::string& name);

}  // namespace hfscf
