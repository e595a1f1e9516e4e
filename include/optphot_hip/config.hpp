// config.hpp -- run configuration and output files with the reference's
// interface (include/optphot/config.hpp): "key = value" files, --key value /
// --key=value command-line overrides with the same keys, defaults and error
// behaviour (std::runtime_error on an unknown key or a bad value), and the
// .npy + meta.json output layout that the reference's Python analysis reads.
// Implemented in gpu_optical_photons_amd/host/{config,io}.cpp
// (liboptphot_host.a).
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "backend.hpp"

namespace optphot {

struct RunOptions {
  std::string backend = "cpu";   // cpu | gpu (this library runs gpu only)
  std::string mode = "records";  // records | tally
  uint64_t n_photons = 1000000;
  int threads = 1;
  int block_size = 256;
  std::string out;  // output directory, "" = none
  TallyConfig tally;
};

// Geometry keys are full lengths (size_x -> hx = size_x / 2); length keys
// accept "inf"; integer keys accept 1e6-style values.
void apply_setting(SimParams& p, RunOptions& o, const std::string& key, const std::string& value);
void load_config_file(const std::string& path, SimParams& p, RunOptions& o);
// --config FILE is applied where it appears; later settings win. Returns
// false when --help / -h was given.
bool parse_command_line(int argc, char** argv, SimParams& p, RunOptions& o);
std::string usage();

std::string surface_name(Surface s);
std::string source_name(SourceType s);
std::string fate_name(Fate f);
std::string params_to_json(const SimParams& p, const RunOptions& o);

// NumPy .npy, format version 1.0, data aligned to 64 bytes.
void write_npy(const std::string& path, const void* data, const std::string& descr,
               const std::vector<std::size_t>& shape, std::size_t item_size);
// <dir>/{fate,n_boundary,n_scatter,t,x,y,z,path}.npy + <dir>/meta.json
void write_records(const std::string& dir, const PhotonRecords& rec, const std::string& meta_json);
// <dir>/{fate_counts,time_hist,xy_hist}.npy + <dir>/meta.json
void write_tally(const std::string& dir, const Tally& tally, const std::string& meta_json);
void write_text_file(const std::string& path, const std::string& text);

}  // namespace optphot
