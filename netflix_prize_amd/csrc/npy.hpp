// Minimal .npy (format 1.0 / 2.0) reader and writer for the rating files
// (data/<set>/*.npy, read by Dataset::load, src/lib.rs:252-278) and the
// prediction outputs ({preds}/<job>.<set>.npy, 1-D or 2-D <f4).
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace npy {

struct Array {
    std::string descr;          // e.g. "<i4", "<i2", "|i1", "<f4"
    std::vector<int64_t> shape;
    std::vector<char> data;
    int64_t size() const {
        int64_t n = 1;
        for (int64_t s : shape) n *= s;
        return n;
    }
};

inline std::string header_value(const std::string& h, const std::string& key) {
    const size_t k = h.find("'" + key + "'");
    if (k == std::string::npos) throw std::runtime_error("npy header lacks " + key);
    size_t p = h.find(':', k);
    if (p == std::string::npos) throw std::runtime_error("bad npy header");
    ++p;
    while (p < h.size() && h[p] == ' ') ++p;
    if (h[p] == '\'') {
        const size_t e = h.find('\'', p + 1);
        return h.substr(p + 1, e - p - 1);
    }
    if (h[p] == '(') {
        const size_t e = h.find(')', p);
        return h.substr(p + 1, e - p - 1);
    }
    size_t e = p;
    while (e < h.size() && h[e] != ',' && h[e] != '}') ++e;
    return h.substr(p, e - p);
}

inline Array read(const std::string& path) {
    FILE* fp = std::fopen(path.c_str(), "rb");
    if (!fp) throw std::runtime_error(path + ": cannot open");
    unsigned char magic[8];
    if (std::fread(magic, 1, 8, fp) != 8 || std::memcmp(magic, "\x93NUMPY", 6) != 0) {
        std::fclose(fp);
        throw std::runtime_error(path + ": not a .npy file");
    }
    uint32_t hlen = 0;
    if (magic[6] == 1) {
        unsigned char b[2];
        if (std::fread(b, 1, 2, fp) != 2) { std::fclose(fp); throw std::runtime_error(path + ": truncated"); }
        hlen = b[0] | (b[1] << 8);
    } else {
        unsigned char b[4];
        if (std::fread(b, 1, 4, fp) != 4) { std::fclose(fp); throw std::runtime_error(path + ": truncated"); }
        hlen = b[0] | (b[1] << 8) | (b[2] << 16) | (static_cast<uint32_t>(b[3]) << 24);
    }
    std::string h(hlen, '\0');
    if (std::fread(&h[0], 1, hlen, fp) != hlen) { std::fclose(fp); throw std::runtime_error(path + ": truncated"); }
    Array a;
    a.descr = header_value(h, "descr");
    if (header_value(h, "fortran_order").find("True") != std::string::npos) {
        std::fclose(fp);
        throw std::runtime_error(path + ": Fortran-order arrays are not supported");
    }
    const std::string shp = header_value(h, "shape");
    size_t p = 0;
    while (p < shp.size()) {
        while (p < shp.size() && (shp[p] == ' ' || shp[p] == ',')) ++p;
        if (p >= shp.size()) break;
        size_t e = p;
        while (e < shp.size() && shp[e] != ',') ++e;
        a.shape.push_back(std::stoll(shp.substr(p, e - p)));
        p = e;
    }
    int itemsize = 0;
    if (a.descr.size() >= 3) itemsize = std::stoi(a.descr.substr(2));
    if (itemsize <= 0 || (a.descr[0] == '>' && itemsize > 1)) {
        std::fclose(fp);
        throw std::runtime_error(path + ": unsupported dtype " + a.descr);
    }
    a.data.resize(static_cast<size_t>(a.size()) * itemsize);
    if (!a.data.empty() && std::fread(a.data.data(), 1, a.data.size(), fp) != a.data.size()) {
        std::fclose(fp);
        throw std::runtime_error(path + ": truncated data");
    }
    std::fclose(fp);
    return a;
}

// Convert to T, widening as the reference does (e.g. i16 item ids -> i32).
template <typename T>
std::vector<T> as(const Array& a, const std::string& path) {
    const char kind = a.descr[1];
    const int sz = std::stoi(a.descr.substr(2));
    std::vector<T> out(static_cast<size_t>(a.size()));
    const char* d = a.data.data();
    for (size_t k = 0; k < out.size(); ++k) {
        if (kind == 'i' && sz == 1) out[k] = static_cast<T>(reinterpret_cast<const int8_t*>(d)[k]);
        else if (kind == 'i' && sz == 2) { int16_t v; std::memcpy(&v, d + 2 * k, 2); out[k] = static_cast<T>(v); }
        else if (kind == 'i' && sz == 4) { int32_t v; std::memcpy(&v, d + 4 * k, 4); out[k] = static_cast<T>(v); }
        else if (kind == 'i' && sz == 8) { int64_t v; std::memcpy(&v, d + 8 * k, 8); out[k] = static_cast<T>(v); }
        else if (kind == 'u' && sz == 1) out[k] = static_cast<T>(reinterpret_cast<const uint8_t*>(d)[k]);
        else if (kind == 'f' && sz == 4) { float v; std::memcpy(&v, d + 4 * k, 4); out[k] = static_cast<T>(v); }
        else if (kind == 'f' && sz == 8) { double v; std::memcpy(&v, d + 8 * k, 8); out[k] = static_cast<T>(v); }
        else throw std::runtime_error(path + ": unsupported dtype " + a.descr);
    }
    return out;
}

template <typename T>
std::vector<T> load(const std::string& path) {
    return as<T>(read(path), path);
}

inline void write_f32(const std::string& path, const float* data, const std::vector<int64_t>& shape) {
    std::string shp = "(";
    for (size_t k = 0; k < shape.size(); ++k) shp += std::to_string(shape[k]) + (shape.size() == 1 ? "," : (k + 1 < shape.size() ? ", " : ""));
    shp += ")";
    std::string h = "{'descr': '<f4', 'fortran_order': False, 'shape': " + shp + ", }";
    const size_t total = 10 + h.size() + 1;
    h.append((64 - total % 64) % 64, ' ');
    h += '\n';
    FILE* fp = std::fopen(path.c_str(), "wb");
    if (!fp) throw std::runtime_error(path + ": cannot create");
    const unsigned char pre[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
    const uint16_t hl = static_cast<uint16_t>(h.size());
    int64_t n = 1;
    for (int64_t s : shape) n *= s;
    bool ok = std::fwrite(pre, 1, 8, fp) == 8 && std::fwrite(&hl, 2, 1, fp) == 1 &&
              std::fwrite(h.data(), 1, h.size(), fp) == h.size() &&
              (n == 0 || std::fwrite(data, sizeof(float), static_cast<size_t>(n), fp) == static_cast<size_t>(n));
    ok = (std::fclose(fp) == 0) && ok;
    if (!ok) throw std::runtime_error(path + ": write failed");
}

}  // namespace npy
