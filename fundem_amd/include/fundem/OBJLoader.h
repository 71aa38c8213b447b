// Wavefront OBJ meshes for TriangleMeshParticle (OBJLoader::loadOBJMesh,
// OBJLoader.h:1-282): same function, arguments and results.
//   - "v x y z" records are vertices, "f a b c ..." records faces; every other
//     record (vn, vt, g, o, usemtl, ...) and '#' comments are skipped;
//   - a face token's vertex index is what precedes its first '/' ("7",
//     "7/2", "7//3", "7/2/3"); 1-based, negative counts back from the last
//     vertex defined so far;
//   - polygons are split into the fan (0, i, i+1);
//   - a relative file name is taken from the executable's folder (argv[0]),
//     else from the working directory.
// On a malformed record, an index of zero or out of range, or a face with
// fewer than three vertices, it reports the line on std::cerr, clears both
// outputs and returns false.
#pragma once
#include <filesystem>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "types.h"

namespace OBJLoader {

inline std::string resolvePath(const std::string& filename, const char* argv0) {
  namespace fs = std::filesystem;
  const fs::path p(filename);
  if (p.is_absolute() || argv0 == nullptr) return p.is_absolute() ? p.string() : (fs::current_path() / p).string();
  const fs::path exe = fs::weakly_canonical(fs::absolute(fs::path(argv0)));
  return (exe.parent_path() / p).string();
}

// Zero-based vertex index of one face token given the vertices read so far.
inline int faceVertex(const std::string& token, int numVertices) {
  const std::string head = token.substr(0, token.find('/'));
  if (head.empty()) throw std::runtime_error("Invalid OBJ face token: " + token);
  const int idx = std::stoi(head);  // leading digits, as std::stoi reads them
  if (idx == 0) throw std::runtime_error("OBJ index cannot be zero.");
  return idx > 0 ? idx - 1 : numVertices + idx;
}

inline bool loadOBJMesh(const std::string& filename, std::vector<double3>& vertices,
                        std::vector<int3>& faces, const int argc, char** argv,
                        const bool verbose = true) {
  vertices.clear();
  faces.clear();
  const std::string path = resolvePath(filename, (argc > 0 && argv) ? argv[0] : nullptr);
  std::ifstream in(path);
  if (!in.is_open()) {
    std::cerr << "[OBJLoader] Failed to open file: " << path << std::endl;
    return false;
  }
  std::string line;
  size_t lineNo = 0;
  try {
    while (std::getline(in, line)) {
      ++lineNo;
      const size_t a = line.find_first_not_of(" \t\r\n");
      if (a == std::string::npos) continue;
      line = line.substr(a, line.find_last_not_of(" \t\r\n") - a + 1);
      if (line[0] == '#') continue;
      if (line.compare(0, 2, "v ") == 0) {
        std::istringstream r(line.substr(2));
        double x, y, z;
        if (!(r >> x >> y >> z))
          throw std::runtime_error("Failed to parse vertex at line " + std::to_string(lineNo));
        vertices.push_back(make_double3(x, y, z));
      } else if (line.compare(0, 2, "f ") == 0) {
        std::istringstream r(line.substr(2));
        std::vector<int> poly;
        std::string tok;
        const int nv = (int)vertices.size();
        while (r >> tok) {
          const int v = faceVertex(tok, nv);
          if (v < 0 || v >= nv)
            throw std::runtime_error("Face index out of range at line " + std::to_string(lineNo));
          poly.push_back(v);
        }
        if (poly.size() < 3)
          throw std::runtime_error("Face has fewer than 3 vertices at line " + std::to_string(lineNo));
        for (size_t i = 1; i + 1 < poly.size(); ++i) faces.push_back(make_int3(poly[0], poly[i], poly[i + 1]));
      }
    }
  } catch (const std::exception& e) {
    std::cerr << "[OBJLoader] Error while reading " << path << " at line " << lineNo << ": " << e.what()
              << std::endl;
    vertices.clear();
    faces.clear();
    return false;
  }
  if (verbose) {
    std::cout << "[OBJLoader] Loaded OBJ: " << path << std::endl;
    std::cout << "  Num vertices = " << vertices.size() << std::endl;
    std::cout << "  Num faces    = " << faces.size() << std::endl;
  }
  return true;
}

}  // namespace OBJLoader
