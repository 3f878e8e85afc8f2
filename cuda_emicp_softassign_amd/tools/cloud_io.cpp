// Point-cloud readers and writers, seeded reduction, scale normalisation and
// (R, t) files for the command-line harness and file tools (cloud_io.h).
#include "cloud_io.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace emsa_io {

namespace {

bool read_file(const std::string& path, std::string& data, std::string& err) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    std::ostringstream ss;
    ss << f.rdbuf();
    data = ss.str();
    return true;
}

// One header line starting at pos (LF or CRLF); pos moves past it.
bool next_line(const std::string& d, size_t& pos, std::string& line) {
    if (pos >= d.size()) return false;
    size_t e = d.find('\n', pos);
    if (e == std::string::npos) e = d.size();
    line = d.substr(pos, e - pos);
    if (!line.empty() && line.back() == '\r') line.pop_back();
    pos = e + 1;
    return true;
}

std::vector<std::string> split(const std::string& s) {
    std::vector<std::string> out;
    std::istringstream is(s);
    std::string w;
    while (is >> w) out.push_back(w);
    return out;
}

std::string lower(std::string s) {
    for (char& c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

bool ends_with(const std::string& s, const std::string& suf) {
    return s.size() >= suf.size() && lower(s.substr(s.size() - suf.size())) == suf;
}

void keep_finite(std::vector<Point>& v) {
    v.erase(std::remove_if(v.begin(), v.end(),
                           [](const Point& p) {
                               return !(std::isfinite(p.x) && std::isfinite(p.y) &&
                                        std::isfinite(p.z));
                           }),
            v.end());
}

// ---- PLY ----------------------------------------------------------------

int ply_type_size(const std::string& t) {
    if (t == "char" || t == "int8" || t == "uchar" || t == "uint8") return 1;
    if (t == "short" || t == "int16" || t == "ushort" || t == "uint16") return 2;
    if (t == "int" || t == "int32" || t == "uint" || t == "uint32" || t == "float" ||
        t == "float32")
        return 4;
    if (t == "double" || t == "float64") return 8;
    return 0;
}

struct PlyProp {
    std::string name, type;
    bool list = false;
    std::string countType;
};

struct PlyElem {
    std::string name;
    size_t count = 0;
    std::vector<PlyProp> props;
};

// Binary scalar of PLY type t at p, as double.
double bin_value(const unsigned char* p, const std::string& t, bool swap) {
    unsigned char b[8];
    const int n = ply_type_size(t);
    for (int i = 0; i < n; ++i) b[i] = swap ? p[n - 1 - i] : p[i];
    if (t == "char" || t == "int8") return (double)(int8_t)b[0];
    if (t == "uchar" || t == "uint8") return (double)b[0];
    if (t == "short" || t == "int16") {
        int16_t v;
        std::memcpy(&v, b, 2);
        return v;
    }
    if (t == "ushort" || t == "uint16") {
        uint16_t v;
        std::memcpy(&v, b, 2);
        return v;
    }
    if (t == "int" || t == "int32") {
        int32_t v;
        std::memcpy(&v, b, 4);
        return v;
    }
    if (t == "uint" || t == "uint32") {
        uint32_t v;
        std::memcpy(&v, b, 4);
        return v;
    }
    if (t == "float" || t == "float32") {
        float v;
        std::memcpy(&v, b, 4);
        return v;
    }
    double v;
    std::memcpy(&v, b, 8);
    return v;
}

bool load_ply(const std::string& d, std::vector<Point>& out, std::string& err) {
    size_t pos = 0;
    std::string line;
    next_line(d, pos, line);  // "ply"
    std::string format;
    std::vector<PlyElem> elems;
    bool ended = false;
    while (next_line(d, pos, line)) {
        const std::vector<std::string> w = split(line);
        if (w.empty()) continue;
        if (w[0] == "format" && w.size() >= 2) {
            format = w[1];
        } else if (w[0] == "element" && w.size() >= 3) {
            PlyElem e;
            e.name = w[1];
            e.count = (size_t)std::strtoull(w[2].c_str(), nullptr, 10);
            elems.push_back(e);
        } else if (w[0] == "property" && !elems.empty()) {
            PlyProp pr;
            if (w.size() >= 5 && w[1] == "list") {
                pr.list = true;
                pr.countType = w[2];
                pr.type = w[3];
                pr.name = w[4];
            } else if (w.size() >= 3) {
                pr.type = w[1];
                pr.name = w[2];
            } else {
                err = "PLY: bad property line";
                return false;
            }
            if (!ply_type_size(pr.type) || (pr.list && !ply_type_size(pr.countType))) {
                err = "PLY: unknown property type in '" + line + "'";
                return false;
            }
            elems.back().props.push_back(pr);
        } else if (w[0] == "end_header") {
            ended = true;
            break;
        }
    }
    if (!ended) return err = "PLY: no end_header", false;
    if (format != "ascii" && format != "binary_little_endian" && format != "binary_big_endian")
        return err = "PLY: unknown format '" + format + "'", false;
    const bool ascii = format == "ascii";
    const bool swap = format == "binary_big_endian";  // hosts here are little-endian
    const char* const txt = d.c_str();
    for (const PlyElem& e : elems) {
        const bool vertex = e.name == "vertex";
        int ix = -1, iy = -1, iz = -1;
        for (size_t k = 0; k < e.props.size(); ++k) {
            if (e.props[k].list) continue;
            if (e.props[k].name == "x") ix = (int)k;
            if (e.props[k].name == "y") iy = (int)k;
            if (e.props[k].name == "z") iz = (int)k;
        }
        if (vertex && (ix < 0 || iy < 0 || iz < 0))
            return err = "PLY: vertex element without x, y, z", false;
        if (vertex) out.reserve(e.count);
        std::vector<double> vals(e.props.size());
        for (size_t i = 0; i < e.count; ++i) {
            for (size_t k = 0; k < e.props.size(); ++k) {
                const PlyProp& pr = e.props[k];
                if (ascii) {
                    char* endp = nullptr;
                    if (pr.list) {
                        const double c = std::strtod(txt + pos, &endp);
                        if (endp == txt + pos) return err = "PLY: short ascii data", false;
                        pos = endp - txt;
                        for (long j = 0; j < (long)c; ++j) {
                            std::strtod(txt + pos, &endp);
                            if (endp == txt + pos) return err = "PLY: short ascii data", false;
                            pos = endp - txt;
                        }
                    } else {
                        vals[k] = std::strtod(txt + pos, &endp);
                        if (endp == txt + pos) return err = "PLY: short ascii data", false;
                        pos = endp - txt;
                    }
                } else {
                    if (pr.list) {
                        const int cs = ply_type_size(pr.countType), is = ply_type_size(pr.type);
                        if (pos + cs > d.size()) return err = "PLY: short binary data", false;
                        const double c =
                            bin_value((const unsigned char*)txt + pos, pr.countType, swap);
                        pos += cs + (size_t)c * is;
                        if (pos > d.size()) return err = "PLY: short binary data", false;
                    } else {
                        const int sz = ply_type_size(pr.type);
                        if (pos + sz > d.size()) return err = "PLY: short binary data", false;
                        vals[k] = bin_value((const unsigned char*)txt + pos, pr.type, swap);
                        pos += sz;
                    }
                }
            }
            if (vertex) out.push_back({(float)vals[ix], (float)vals[iy], (float)vals[iz]});
        }
        if (vertex) return true;  // nothing after the vertices is needed
    }
    return err = "PLY: no vertex element", false;
}

// ---- PCD ----------------------------------------------------------------

bool load_pcd(const std::string& d, std::vector<Point>& out, std::string& err) {
    size_t pos = 0;
    std::string line, data;
    std::vector<std::string> fields, types;
    std::vector<int> sizes, counts;
    size_t points = 0;
    while (next_line(d, pos, line)) {
        const std::vector<std::string> w = split(line);
        if (w.empty() || w[0][0] == '#') continue;
        if (w[0] == "FIELDS") fields.assign(w.begin() + 1, w.end());
        if (w[0] == "SIZE")
            for (size_t i = 1; i < w.size(); ++i) sizes.push_back(std::atoi(w[i].c_str()));
        if (w[0] == "TYPE") types.assign(w.begin() + 1, w.end());
        if (w[0] == "COUNT")
            for (size_t i = 1; i < w.size(); ++i) counts.push_back(std::atoi(w[i].c_str()));
        if (w[0] == "POINTS" && w.size() > 1) points = (size_t)std::strtoull(w[1].c_str(), 0, 10);
        if (w[0] == "DATA") {
            data = w.size() > 1 ? w[1] : "";
            break;
        }
    }
    if (counts.empty()) counts.assign(fields.size(), 1);
    if (fields.empty() || sizes.size() != fields.size() || types.size() != fields.size() ||
        counts.size() != fields.size())
        return err = "PCD: bad header", false;
    int fx = -1, fy = -1, fz = -1;
    std::vector<int> col(fields.size()), off(fields.size());
    int ncol = 0, stride = 0;
    for (size_t i = 0; i < fields.size(); ++i) {
        col[i] = ncol;
        off[i] = stride;
        ncol += counts[i];
        stride += sizes[i] * counts[i];
        if (fields[i] == "x") fx = (int)i;
        if (fields[i] == "y") fy = (int)i;
        if (fields[i] == "z") fz = (int)i;
    }
    if (fx < 0 || fy < 0 || fz < 0) return err = "PCD: no x, y, z fields", false;
    out.reserve(points);
    if (data == "ascii") {
        std::vector<double> v(ncol);
        std::string l;
        while (out.size() < points && next_line(d, pos, l)) {
            const char* s = l.c_str();
            int k = 0;
            for (; k < ncol; ++k) {
                char* e = nullptr;
                v[k] = std::strtod(s, &e);
                if (e == s) break;
                s = e;
            }
            if (k == 0) continue;
            if (k < ncol) return err = "PCD: short ascii row", false;
            out.push_back({(float)v[col[fx]], (float)v[col[fy]], (float)v[col[fz]]});
        }
        if (out.size() < points) return err = "PCD: fewer rows than POINTS", false;
        return true;
    }
    if (data == "binary") {
        if (pos + points * (size_t)stride > d.size()) return err = "PCD: short binary data", false;
        auto get = [&](size_t base, int f) -> float {
            const unsigned char* p = (const unsigned char*)d.data() + base + off[f];
            if (types[f] == "F" && sizes[f] == 4) {
                float x;
                std::memcpy(&x, p, 4);
                return x;
            }
            if (types[f] == "F" && sizes[f] == 8) {
                double x;
                std::memcpy(&x, p, 8);
                return (float)x;
            }
            return std::nanf("");
        };
        for (size_t i = 0; i < points; ++i) {
            const size_t b = pos + i * stride;
            out.push_back({get(b, fx), get(b, fy), get(b, fz)});
        }
        return true;
    }
    return err = "PCD: DATA '" + data + "' not supported (ascii, binary)", false;
}

// ---- text ---------------------------------------------------------------

bool load_text(const std::string& d, std::vector<Point>& out, std::string&) {
    size_t pos = 0;
    std::string l;
    while (next_line(d, pos, l)) {
        const char* s = l.c_str();
        while (*s == ' ' || *s == '\t') ++s;
        if (*s == '#' || *s == 0) continue;
        double v[3];
        int k = 0;
        for (; k < 3; ++k) {
            char* e = nullptr;
            v[k] = std::strtod(s, &e);
            if (e == s) break;
            s = e;
        }
        if (k == 3) out.push_back({(float)v[0], (float)v[1], (float)v[2]});
    }
    return true;
}

uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

bool write_text(const std::string& path, const std::string& body, std::string& err) {
    std::ofstream f(path, std::ios::binary);
    if (!f) return err = "cannot write " + path, false;
    f << body;
    if (!f) return err = "write failed: " + path, false;
    return true;
}

std::string rows(const std::vector<Point>& c) {
    std::string s;
    char buf[96];
    for (const Point& p : c) {
        std::snprintf(buf, sizeof(buf), "%.9g %.9g %.9g\n", p.x, p.y, p.z);
        s += buf;
    }
    return s;
}

}  // namespace

bool load_cloud(const std::string& path, std::vector<Point>& out, std::string& err) {
    out.clear();
    std::string d;
    if (!read_file(path, d, err)) return false;
    bool ok;
    if (d.compare(0, 3, "ply") == 0)
        ok = load_ply(d, out, err);
    else if (ends_with(path, ".pcd"))
        ok = load_pcd(d, out, err);
    else
        ok = load_text(d, out, err);
    if (!ok) {
        err = path + ": " + err;
        return false;
    }
    keep_finite(out);
    return true;
}

bool save_pcd_ascii(const std::string& path, const std::vector<Point>& c, std::string& err) {
    char h[512];
    std::snprintf(h, sizeof(h),
                  "# .PCD v0.7 - Point Cloud Data file format\nVERSION 0.7\nFIELDS x y z\n"
                  "SIZE 4 4 4\nTYPE F F F\nCOUNT 1 1 1\nWIDTH %zu\nHEIGHT 1\n"
                  "VIEWPOINT 0 0 0 1 0 0 0\nPOINTS %zu\nDATA ascii\n",
                  c.size(), c.size());
    return write_text(path, h + rows(c), err);
}

bool save_ply_ascii(const std::string& path, const std::vector<Point>& c, std::string& err) {
    char h[256];
    std::snprintf(h, sizeof(h),
                  "ply\nformat ascii 1.0\nelement vertex %zu\nproperty float x\n"
                  "property float y\nproperty float z\nend_header\n",
                  c.size());
    return write_text(path, h + rows(c), err);
}

bool save_xyz(const std::string& path, const std::vector<Point>& c, std::string& err) {
    return write_text(path, rows(c), err);
}

void points_reduction(std::vector<Point>& cloud, float percent, uint64_t seed) {
    const double keep = percent / 100.0;
    uint64_t s = seed;
    std::vector<Point> out;
    out.reserve(cloud.size());
    for (const Point& p : cloud) {
        const double u = (double)(splitmix64(s) >> 11) * 0x1.0p-53;
        if (u < keep) out.push_back(p);
    }
    cloud.swap(out);
}

double covariance_trace(const std::vector<Point>& c) {
    if (c.empty()) return 0.0;
    double m[3] = {0, 0, 0};
    for (const Point& p : c) {
        m[0] += p.x;
        m[1] += p.y;
        m[2] += p.z;
    }
    for (double& v : m) v /= (double)c.size();
    double tr = 0;
    for (const Point& p : c) {
        const double a = p.x - m[0], b = p.y - m[1], d = p.z - m[2];
        tr += a * a + b * b + d * d;
    }
    return tr / (double)c.size();
}

void align_scale_once(std::vector<Point>& target, std::vector<Point>& source, double tr[4]) {
    tr[0] = covariance_trace(target);
    tr[1] = covariance_trace(source);
    for (int k = 0; k < 2; ++k) {
        std::vector<Point>& c = k == 0 ? target : source;
        const float s = (float)(std::sqrt(0.003) / std::sqrt(tr[k]));
        for (Point& p : c) {
            p.x *= s;
            p.y *= s;
            p.z *= s;
        }
    }
    tr[2] = covariance_trace(target);
    tr[3] = covariance_trace(source);
}

bool save_rt(const std::string& path, const float R[9], const float t[3]) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    for (int r = 0; r < 9; ++r) {
        std::fprintf(f, "%f ", R[r]);
        if (r % 3 == 2) std::fprintf(f, "\n");
    }
    for (int r = 0; r < 3; ++r) std::fprintf(f, "%f ", t[r]);
    std::fprintf(f, "\n");
    return std::fclose(f) == 0;
}

bool load_rt(const std::string& path, float R[9], float t[3]) {
    FILE* f = std::fopen(path.c_str(), "r");
    if (!f) return false;
    float v[12];
    int n = 0;
    while (n < 12 && std::fscanf(f, "%f", &v[n]) == 1) ++n;
    std::fclose(f);
    if (n != 12) return false;
    std::memcpy(R, v, 9 * sizeof(float));
    std::memcpy(t, v + 9, 3 * sizeof(float));
    return true;
}

void transform(std::vector<Point>& c, const float R[9], const float t[3]) {
    for (Point& p : c) {
        const float x = p.x, y = p.y, z = p.z;
        p.x = R[0] * x + R[1] * y + R[2] * z + t[0];
        p.y = R[3] * x + R[4] * y + R[5] * z + t[1];
        p.z = R[6] * x + R[7] * y + R[8] * z + t[2];
    }
}

}  // namespace emsa_io
