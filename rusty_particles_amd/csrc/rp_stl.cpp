// rp_stl.cpp — see rp_stl.hpp.
#include "rp_stl.hpp"

#include <cctype>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

#include "../../include/rp_abi.h"

namespace {

bool parse_ascii(const std::string& text, std::vector<float>& v9, std::vector<float>& n3, std::string& err) {
  std::istringstream in(text);
  std::string tok;
  in >> tok;  // "solid"
  std::string line;
  std::getline(in, line);  // rest of the solid line (name)
  size_t ntri = 0;
  while (in >> tok) {
    if (tok == "endsolid") break;
    if (tok != "facet") {
      err = "ASCII STL: expected 'facet', got '" + tok + "'";
      return false;
    }
    std::string kw;
    float n[3];
    if (!(in >> kw) || kw != "normal" || !(in >> n[0] >> n[1] >> n[2])) {
      err = "ASCII STL: bad facet normal";
      return false;
    }
    std::string a, b;
    if (!(in >> a >> b) || a != "outer" || b != "loop") {
      err = "ASCII STL: expected 'outer loop'";
      return false;
    }
    float v[9];
    for (int k = 0; k < 3; ++k) {
      if (!(in >> kw) || kw != "vertex" || !(in >> v[3 * k] >> v[3 * k + 1] >> v[3 * k + 2])) {
        err = "ASCII STL: bad vertex";
        return false;
      }
    }
    if (!(in >> a) || a != "endloop" || !(in >> b) || b != "endfacet") {
      err = "ASCII STL: expected 'endloop endfacet'";
      return false;
    }
    v9.insert(v9.end(), v, v + 9);
    n3.insert(n3.end(), n, n + 3);
    ++ntri;
  }
  (void)ntri;
  return true;
}

bool looks_ascii(const std::string& data) {
  size_t p = 0;
  while (p < data.size() && std::isspace((unsigned char)data[p])) ++p;
  if (data.compare(p, 5, "solid") != 0) return false;
  // A binary file may also begin with "solid" in its header: require the
  // next keyword to be "facet" or "endsolid".
  size_t eol = data.find('\n', p);
  if (eol == std::string::npos) return false;
  size_t q = eol + 1;
  while (q < data.size() && std::isspace((unsigned char)data[q])) ++q;
  return data.compare(q, 5, "facet") == 0 || data.compare(q, 8, "endsolid") == 0;
}

}  // namespace

bool rp_stl_load(const std::string& path, std::vector<float>& v9, std::vector<float>& n3, std::string& err) {
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    err = "cannot open STL file: " + path;
    return false;
  }
  std::string data((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  if (looks_ascii(data)) return parse_ascii(data, v9, n3, err);
  if (data.size() < 84) {
    err = "binary STL shorter than its 84-byte header: " + path;
    return false;
  }
  uint32_t n;
  std::memcpy(&n, data.data() + 80, 4);
  if (data.size() < 84 + (size_t)n * 50) {
    err = "binary STL truncated: " + path;
    return false;
  }
  const char* p = data.data() + 84;
  for (uint32_t t = 0; t < n; ++t, p += 50) {
    float rec[12];
    std::memcpy(rec, p, 48);
    n3.insert(n3.end(), rec, rec + 3);
    v9.insert(v9.end(), rec + 3, rec + 12);
  }
  return true;
}

extern "C" rp_status rp_stl_read(const char* path, size_t* ntri, float* v9, float* n3) {
  if (!path || !ntri) return RP_ERR_INVALID;
  std::vector<float> vv, nn;
  std::string err;
  if (!rp_stl_load(path, vv, nn, err)) return RP_ERR_IO;
  size_t n = nn.size() / 3;
  if (v9 && n3) {
    if (*ntri < n) return RP_ERR_INVALID;
    std::memcpy(v9, vv.data(), vv.size() * sizeof(float));
    std::memcpy(n3, nn.data(), nn.size() * sizeof(float));
  }
  *ntri = n;
  return RP_OK;
}
