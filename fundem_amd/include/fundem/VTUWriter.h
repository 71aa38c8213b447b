// XML UnstructuredGrid (.vtu) files with inline binary arrays, the on-disk
// format of the reference's output frames (LSParticle.h:426-756,
// SolidInteraction.h:153-316, VBondedInteraction.h:262-448): little-endian,
// header_type UInt64, every binary DataArray the base64 text of
// [UInt64 byte count][values], a FieldData block with TIME (Float32) and
// STEP (Int32) in ASCII, and one Piece. Array names and types are the
// reference's, so post-processing scripts read either solver's frames.
#pragma once
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <vector>

namespace fundem {
namespace vtu {

// RFC 4648 base64 of the UInt64 byte count followed by the bytes.
inline void appendBase64(std::string& out, const void* data, size_t nBytes) {
  static const char kAlphabet[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
  const uint64_t count = nBytes;
  const size_t total = sizeof count + nBytes;
  auto byteAt = [&](size_t i) -> uint32_t {
    if (i < sizeof count) return reinterpret_cast<const uint8_t*>(&count)[i];
    return static_cast<const uint8_t*>(data)[i - sizeof count];
  };
  out.reserve(out.size() + 4 * ((total + 2) / 3));
  for (size_t i = 0; i < total; i += 3) {
    const size_t left = total - i;
    const uint32_t w = (byteAt(i) << 16) | (left > 1 ? byteAt(i + 1) << 8 : 0u) | (left > 2 ? byteAt(i + 2) : 0u);
    out += kAlphabet[(w >> 18) & 63];
    out += kAlphabet[(w >> 12) & 63];
    out += left > 1 ? kAlphabet[(w >> 6) & 63] : '=';
    out += left > 2 ? kAlphabet[w & 63] : '=';
  }
}

template <class T>
struct TypeName;
template <>
struct TypeName<float> {
  static constexpr const char* value = "Float32";
};
template <>
struct TypeName<int32_t> {
  static constexpr const char* value = "Int32";
};
template <>
struct TypeName<uint8_t> {
  static constexpr const char* value = "UInt8";
};

// mkdir -p
inline void makeDirs(const std::string& path) {
  std::string part;
  std::stringstream ss(path);
  std::string item;
  if (!path.empty() && path[0] == '/') part = "/";
  while (std::getline(ss, item, '/')) {
    if (item.empty()) continue;
    part += item;
    ::mkdir(part.c_str(), 0755);
    part += '/';
  }
  struct stat st;
  if (::stat(path.c_str(), &st) != 0 || !S_ISDIR(st.st_mode))
    throw std::runtime_error("Cannot create directory: " + path);
}

// dir/<stem>_<frame, 4 digits>.vtu
inline std::string framePath(const std::string& dir, const char* stem, size_t frame) {
  std::ostringstream s;
  s << dir << '/' << stem << '_' << std::setw(4) << std::setfill('0') << frame << ".vtu";
  return s.str();
}

// One Piece of points and (optional) cells with point arrays, assembled in
// memory and written in one go.
class Piece {
 public:
  Piece(double time, size_t step, size_t numPoints, size_t numCells) : points_(numPoints), cells_(numCells) {
    std::ostringstream h;
    h << "<?xml version=\"1.0\"?>\n"
      << "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\" "
         "header_type=\"UInt64\">\n"
      << "<UnstructuredGrid>\n"
      << "<FieldData>\n"
      << "<DataArray type=\"Float32\" Name=\"TIME\" NumberOfTuples=\"1\" format=\"ascii\">"
      << static_cast<float>(time) << "</DataArray>\n"
      << "<DataArray type=\"Int32\" Name=\"STEP\" NumberOfTuples=\"1\" format=\"ascii\">"
      << static_cast<int32_t>(step) << "</DataArray>\n"
      << "</FieldData>\n"
      << "<Piece NumberOfPoints=\"" << numPoints << "\" NumberOfCells=\"" << numCells << "\">\n";
    text_ = h.str();
  }

  void points(const std::vector<float>& xyz) {
    check(xyz.size(), 3 * points_, "Points");
    text_ += "<Points>\n";
    array("", xyz, 3, false);
    text_ += "</Points>\n";
  }
  // Cells with explicit connectivity, or none (empty ASCII arrays).
  void cells(const std::vector<int32_t>& connectivity, const std::vector<int32_t>& offsets,
             const std::vector<uint8_t>& types) {
    check(offsets.size(), cells_, "offsets");
    check(types.size(), cells_, "types");
    text_ += "<Cells>\n";
    array("connectivity", connectivity, 1, false);
    array("offsets", offsets, 1, false);
    array("types", types, 1, false);
    text_ += "</Cells>\n";
  }
  void noCells() {
    text_ += "<Cells>\n"
             "<DataArray type=\"Int32\" Name=\"connectivity\" format=\"ascii\"></DataArray>\n"
             "<DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\"></DataArray>\n"
             "<DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\"></DataArray>\n"
             "</Cells>\n";
  }
  // Opens PointData; `vectors` names the active vector array (may be empty).
  void beginPointData(const char* vectors) {
    text_ += vectors && *vectors ? std::string("<PointData Vectors=\"") + vectors + "\">\n" : "<PointData>\n";
  }
  // Opens PointData with `scalars` as the active scalar array.
  void beginPointDataScalars(const char* scalars) {
    text_ += std::string("<PointData Scalars=\"") + scalars + "\">\n";
  }
  template <class T>
  void pointArray(const char* name, const std::vector<T>& values, int components) {
    check(values.size(), points_ * components, name);
    array(name, values, components, true);
  }
  void write(const std::string& path) {
    text_ += "</PointData>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n";
    std::ofstream out(path, std::ios::binary);
    if (!out) throw std::runtime_error("Cannot open for writing: " + path);
    out.write(text_.data(), (std::streamsize)text_.size());
    if (!out) throw std::runtime_error("Write error on: " + path);
  }

 private:
  static void check(size_t got, size_t want, const char* what) {
    if (got != want)
      throw std::logic_error(std::string("vtu: array ") + what + " has " + std::to_string(got) +
                             " values, expected " + std::to_string(want));
  }
  template <class T>
  void array(const char* name, const std::vector<T>& values, int components, bool pointData) {
    text_ += "<DataArray type=\"";
    text_ += TypeName<T>::value;
    text_ += '"';
    if (*name) {
      text_ += " Name=\"";
      text_ += name;
      text_ += '"';
    }
    if (components > 1 || (!pointData && !*name)) {
      text_ += " NumberOfComponents=\"" + std::to_string(components) + '"';
    }
    text_ += " format=\"binary\">\n";
    appendBase64(text_, values.data(), values.size() * sizeof(T));
    text_ += "\n</DataArray>\n";
  }

  size_t points_, cells_;
  std::string text_;
};

}  // namespace vtu
}  // namespace fundem
