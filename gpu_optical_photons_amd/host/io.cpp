// Output files read by the reference's Python analysis: NumPy .npy version
// 1.0 arrays and meta.json (file set and dtypes of src/core/io.cpp:14-74).
// Arrays are written in the host's byte order and declared little-endian;
// x86-64 hosts of MI355X are little-endian.
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <stdexcept>

#include "optphot_hip/config.hpp"

namespace optphot {

void write_npy(const std::string& path, const void* data, const std::string& descr,
               const std::vector<std::size_t>& shape, std::size_t item_size) {
  std::size_t count = 1;
  std::string dims;
  for (std::size_t i = 0; i < shape.size(); ++i) {
    dims += (i ? ", " : "") + std::to_string(shape[i]);
    count *= shape[i];
  }
  if (shape.size() == 1) dims += ",";  // Python's one-element tuple
  std::string header =
      "{'descr': '" + descr + "', 'fortran_order': False, 'shape': (" + dims + "), }";
  // 10-byte preamble + header + '\n', padded with spaces so that the array
  // data starts on a 64-byte boundary.
  const std::size_t used = 10 + header.size() + 1;
  header.append((64 - used % 64) % 64, ' ');
  header += '\n';
  if (header.size() > 0xffff) throw std::runtime_error("npy header too long for '" + path + "'");

  std::ofstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot write '" + path + "'");
  const unsigned char pre[10] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0,
                                 static_cast<unsigned char>(header.size() & 0xff),
                                 static_cast<unsigned char>(header.size() >> 8)};
  f.write(reinterpret_cast<const char*>(pre), sizeof(pre));
  f.write(header.data(), static_cast<std::streamsize>(header.size()));
  f.write(static_cast<const char*>(data), static_cast<std::streamsize>(count * item_size));
  if (!f) throw std::runtime_error("cannot write '" + path + "'");
}

void write_text_file(const std::string& path, const std::string& text) {
  std::FILE* f = std::fopen(path.c_str(), "w");
  bool ok = f != nullptr;
  if (ok) ok = std::fputs(text.c_str(), f) >= 0 && std::fputc('\n', f) != EOF;
  if (f && std::fclose(f) != 0) ok = false;
  if (!ok) throw std::runtime_error("cannot write '" + path + "'");
}

void write_records(const std::string& dir, const PhotonRecords& r, const std::string& meta) {
  std::filesystem::create_directories(dir);
  const std::vector<std::size_t> n{r.size()};
  write_npy(dir + "/fate.npy", r.fate.data(), "|u1", n, 1);
  write_npy(dir + "/n_boundary.npy", r.n_boundary.data(), "<u4", n, 4);
  write_npy(dir + "/n_scatter.npy", r.n_scatter.data(), "<u4", n, 4);
  const std::pair<const char*, const std::vector<float>*> floats[] = {
      {"t", &r.t}, {"x", &r.x}, {"y", &r.y}, {"z", &r.z}, {"path", &r.path}};
  for (const auto& a : floats) write_npy(dir + "/" + a.first + ".npy", a.second->data(), "<f4", n, 4);
  write_text_file(dir + "/meta.json", meta);
}

void write_tally(const std::string& dir, const Tally& t, const std::string& meta) {
  // the hit map is square: xy_bins x xy_bins words, row-major [iy][ix]
  const std::size_t words = t.xy_hist.size();
  const auto side = static_cast<std::size_t>(std::llround(std::sqrt(static_cast<double>(words))));
  if (side * side != words)
    throw std::runtime_error("hit map of " + std::to_string(words) + " words is not square");
  struct Array {
    const char* file;
    const uint64_t* data;
    std::vector<std::size_t> shape;
  };
  const Array arrays[] = {{"fate_counts", t.fate_counts.data(), {t.fate_counts.size()}},
                          {"time_hist", t.time_hist.data(), {t.time_hist.size()}},
                          {"xy_hist", t.xy_hist.data(), {side, side}}};
  std::filesystem::create_directories(dir);
  for (const Array& a : arrays)
    write_npy(dir + "/" + a.file + ".npy", a.data, "<u8", a.shape, sizeof(uint64_t));
  write_text_file(dir + "/meta.json", meta);
}

}  // namespace optphot
