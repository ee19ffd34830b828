// Config (.cfg), model (.mod) and XYZ readers / writer of the BD driver.
//
// Formats and key semantics follow MoleculeIO (src/util/molecule_io.h:22-105,
// src/util/molecule_io.cc:25-414):
//  - config: "key value" lines, '#' comments, duplicate keys are an error;
//    GetFloatKey falls back to the default outside the open range (min, max),
//    GetIntKey outside the closed range [min, max];
//  - model: "key rest-of-line" lines (par / steric / bond ...), repeatable;
//  - XYZ: npos, then "step time Lx Ly Lz", then "type x y z" lines; radii
//    come from the model's "par type radius" lines (1.0 if the type is not
//    listed). Reading stops at the first empty line.
// The writer keeps particle types and writes the comment line with full
// precision plus the Gaussian-stream seed as a sixth field (ignored by the
// reference reader, which scans five), so that a frame is a restart file.
#pragma once
#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

namespace sdt {

class MoleculeIO {
  public:
    // false with a message in err() on failure
    bool parse_config(const char* path);
    bool parse_model(const char* path);
    bool parse_xyz(const char* path);

    double get_float_key(const char* key, double min_value, double max_value,
                         double default_value) const;
    int get_int_key(const char* key, int min_value, int max_value, int default_value) const;
    const char* get_string_key(const char* key) const;   // nullptr if absent
    int line_count(const char* key) const;
    const char* get_line(const char* key, int idx) const;

    // particles of the last parse_xyz
    int npos() const { return (int)rdi_.size(); }
    const std::vector<double>& pos() const { return pos_; }
    const std::vector<double>& rdi() const { return rdi_; }
    const std::vector<std::string>& types() const { return types_; }
    double box[3] = {0.0, 0.0, 0.0};
    long long start_step = 0;
    double start_time = 0.0;
    bool has_seed = false;
    uint64_t seed = 0;

    const std::string& err() const { return err_; }
    const std::vector<std::string>& warnings() const { return warn_; }

  private:
    std::map<std::string, std::string> keys_;
    std::multimap<std::string, std::string> lines_;
    std::vector<double> pos_, rdi_;
    std::vector<std::string> types_;
    std::string err_;
    mutable std::vector<std::string> warn_;
};

// One XYZ frame: npos, "step time Lx Ly Lz seed", "type x y z" (%.17g).
bool write_xyz_frame(FILE* fp, int npos, const double box[3], long long step, double time,
                     uint64_t seed, const std::vector<std::string>& types, const double* pos);

}  // namespace sdt
