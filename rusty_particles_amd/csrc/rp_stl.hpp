// rp_stl.hpp — STL reader with the semantics of Mesh::load_stl
// (src/mesh.rs:125-143), which reads through stl_io 0.7.0 (Cargo.lock):
// binary (80-byte header, u32 count, 50-byte records) or ASCII
// ("solid ... facet normal ... outer loop / vertex x3 ... endsolid").
// Triangles keep file order; each carries its three vertices and the facet
// normal stored in the file.  stl_io de-duplicates vertices into an indexed
// mesh, which does not change any coordinate, so the flat triangle list is
// equivalent.
#pragma once

#include <string>
#include <vector>

// Appends ntri*9 vertex floats to v9 and ntri*3 normal floats to n3.
// Returns false and sets err when the file cannot be read or parsed.
bool rp_stl_load(const std::string& path, std::vector<float>& v9, std::vector<float>& n3, std::string& err);
