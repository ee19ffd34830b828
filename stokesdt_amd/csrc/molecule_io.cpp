// Config / model / XYZ parsing for the BD driver (formats of
// src/util/molecule_io.cc:25-414; see molecule_io.hpp).
#include "molecule_io.hpp"
#include <cerrno>
#include <cinttypes>
#include <cstdlib>
#include <cstring>

namespace sdt {

namespace {

// splits on spaces and tabs
std::vector<std::string> tokens(const std::string& line) {
    std::vector<std::string> out;
    size_t i = 0;
    while (i < line.size()) {
        while (i < line.size() && (line[i] == ' ' || line[i] == '\t')) i++;
        size_t j = i;
        while (j < line.size() && line[j] != ' ' && line[j] != '\t') j++;
        if (j > i) out.push_back(line.substr(i, j - i));
        i = j;
    }
    return out;
}

bool read_lines(const char* path, std::vector<std::string>* lines, std::string* err) {
    FILE* fp = fopen(path, "r");
    if (!fp) {
        *err = std::string("cannot open ") + path + ": " + strerror(errno);
        return false;
    }
    std::string cur;
    int c;
    while ((c = fgetc(fp)) != EOF) {
        if (c == '\n') {
            lines->push_back(cur);
            cur.clear();
        } else if (c != '\r') {
            cur.push_back((char)c);
        }
    }
    if (!cur.empty()) lines->push_back(cur);
    fclose(fp);
    return true;
}

bool is_comment(const std::vector<std::string>& tk) { return !tk.empty() && tk[0][0] == '#'; }

}  // namespace

bool MoleculeIO::parse_config(const char* path) {
    // src/util/molecule_io.cc:25-68: key = first token, value = second
    std::vector<std::string> lines;
    if (!read_lines(path, &lines, &err_)) return false;
    keys_.clear();
    for (const std::string& l : lines) {
        const std::vector<std::string> tk = tokens(l);
        if (tk.size() < 2 || is_comment(tk)) continue;
        if (keys_.count(tk[0])) {
            err_ = "duplicate key \"" + tk[0] + "\" in " + path;
            return false;
        }
        keys_[tk[0]] = tk[1];
    }
    return true;
}

bool MoleculeIO::parse_model(const char* path) {
    // src/util/molecule_io.cc:71-110: key = first token, value = the rest
    std::vector<std::string> lines;
    if (!read_lines(path, &lines, &err_)) return false;
    lines_.clear();
    for (const std::string& l : lines) {
        const std::vector<std::string> tk = tokens(l);
        if (tk.size() < 2 || is_comment(tk)) continue;
        const size_t k = l.find(tk[0]) + tk[0].size();
        size_t v = l.find_first_not_of(" \t", k);
        lines_.insert({tk[0], v == std::string::npos ? std::string() : l.substr(v)});
    }
    return true;
}

bool MoleculeIO::parse_xyz(const char* path) {
    // src/util/molecule_io.cc:113-207
    std::map<std::string, double> par;
    for (int i = 0; i < line_count("par"); i++) {
        const std::vector<std::string> tk = tokens(get_line("par", i));
        if (tk.size() >= 2) par[tk[0]] = atof(tk[1].c_str());
    }
    std::vector<std::string> lines;
    if (!read_lines(path, &lines, &err_)) return false;
    if (lines.size() < 2) {
        err_ = std::string("invalid XYZ file ") + path;
        return false;
    }
    const int n = atoi(lines[0].c_str());
    if (n <= 0) {
        err_ = std::string("invalid particle count in ") + path;
        return false;
    }
    {
        const std::vector<std::string> tk = tokens(lines[1]);
        if (tk.size() < 5) {
            err_ = std::string("invalid XYZ comment line in ") + path;
            return false;
        }
        start_step = atoll(tk[0].c_str());
        start_time = atof(tk[1].c_str());
        for (int d = 0; d < 3; d++) box[d] = atof(tk[2 + d].c_str());
        has_seed = false;
        if (tk.size() >= 6) {
            char* end = nullptr;
            const unsigned long long sv = strtoull(tk[5].c_str(), &end, 10);
            if (end && *end == '\0') {
                has_seed = true;
                seed = (uint64_t)sv;
            }
        }
        if (start_step < 0 || start_time < 0.0 || box[0] <= 0.0 || box[1] <= 0.0 ||
            box[2] <= 0.0) {
            err_ = std::string("invalid XYZ comment line in ") + path;
            return false;
        }
    }
    pos_.clear();
    rdi_.clear();
    types_.clear();
    for (size_t i = 2; i < lines.size(); i++) {
        if (lines[i].empty()) break;   // the reference stops at an empty line
        const std::vector<std::string> tk = tokens(lines[i]);
        if (tk.empty()) continue;
        if (tk.size() < 4) {
            err_ = std::string("invalid particle line in ") + path;
            return false;
        }
        types_.push_back(tk[0]);
        for (int d = 0; d < 3; d++) pos_.push_back(atof(tk[1 + d].c_str()));
        auto it = par.find(tk[0]);
        rdi_.push_back(it == par.end() ? 1.0 : it->second);
    }
    if ((int)rdi_.size() != n) {
        err_ = std::string("particle count does not match the header in ") + path;
        return false;
    }
    return true;
}

double MoleculeIO::get_float_key(const char* key, double min_value, double max_value,
                                 double default_value) const {
    // src/util/molecule_io.cc:228-249 (open range)
    auto it = keys_.find(key);
    if (it == keys_.end()) return default_value;
    const double v = atof(it->second.c_str());
    if (v >= max_value || v <= min_value) {
        warn_.push_back(std::string("key \"") + key + "\" out of range, using the default");
        return default_value;
    }
    return v;
}

int MoleculeIO::get_int_key(const char* key, int min_value, int max_value,
                            int default_value) const {
    // src/util/molecule_io.cc:252-273 (closed range)
    auto it = keys_.find(key);
    if (it == keys_.end()) return default_value;
    const int v = atoi(it->second.c_str());
    if (v > max_value || v < min_value) {
        warn_.push_back(std::string("key \"") + key + "\" out of range, using the default");
        return default_value;
    }
    return v;
}

const char* MoleculeIO::get_string_key(const char* key) const {
    auto it = keys_.find(key);
    return it == keys_.end() ? nullptr : it->second.c_str();
}

int MoleculeIO::line_count(const char* key) const { return (int)lines_.count(key); }

const char* MoleculeIO::get_line(const char* key, int idx) const {
    auto r = lines_.equal_range(key);
    auto it = r.first;
    for (int i = 0; i < idx && it != r.second; i++) ++it;
    return it == r.second ? nullptr : it->second.c_str();
}

bool write_xyz_frame(FILE* fp, int npos, const double box[3], long long step, double time,
                     uint64_t seed, const std::vector<std::string>& types, const double* pos) {
    if (fprintf(fp, "%d\n%lld %.17g %.17g %.17g %.17g %" PRIu64 "\n", npos, step, time, box[0],
                box[1], box[2], seed) < 0)
        return false;
    for (int i = 0; i < npos; i++) {
        const char* t = (i < (int)types.size()) ? types[i].c_str() : "X";
        if (fprintf(fp, "%s %.17g %.17g %.17g\n", t, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]) <
            0)
            return false;
    }
    return true;
}

}  // namespace sdt
