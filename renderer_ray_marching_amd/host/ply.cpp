// PLY reader / writer. Reads what Pointcloud::load_ply (src/pointcloud.cpp:43-58) reads through
// happly: the "vertex" element's x, y, z as double (float or double properties) and red, green,
// blue as unsigned char (uchar properties only), in ascii, binary_little_endian or
// binary_big_endian files. Other elements and properties (faces, normals, lists) are skipped.
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "rm_io.h"

namespace rmh {

namespace {

enum class PType { I8, U8, I16, U16, I32, U32, F32, F64 };

int type_size(PType t) {
    switch (t) {
        case PType::I8: case PType::U8: return 1;
        case PType::I16: case PType::U16: return 2;
        case PType::I32: case PType::U32: case PType::F32: return 4;
        case PType::F64: return 8;
    }
    return 0;
}

bool parse_type(const std::string& s, PType& t) {
    if (s == "char" || s == "int8") t = PType::I8;
    else if (s == "uchar" || s == "uint8") t = PType::U8;
    else if (s == "short" || s == "int16") t = PType::I16;
    else if (s == "ushort" || s == "uint16") t = PType::U16;
    else if (s == "int" || s == "int32") t = PType::I32;
    else if (s == "uint" || s == "uint32") t = PType::U32;
    else if (s == "float" || s == "float32") t = PType::F32;
    else if (s == "double" || s == "float64") t = PType::F64;
    else return false;
    return true;
}

struct Prop {
    std::string name;
    bool is_list = false;
    PType type = PType::F32;
    PType count_type = PType::U8;
};

struct Elem {
    std::string name;
    int64_t count = 0;
    std::vector<Prop> props;
};

enum class Format { Ascii, BinLE, BinBE };

// Value of a binary scalar at p (file byte order `swap` = big endian) as double.
double read_scalar(const unsigned char* p, PType t, bool swap) {
    unsigned char b[8];
    const int sz = type_size(t);
    for (int k = 0; k < sz; ++k) b[k] = swap ? p[sz - 1 - k] : p[k];
    switch (t) {
        case PType::I8: { int8_t v; std::memcpy(&v, b, 1); return v; }
        case PType::U8: { uint8_t v; std::memcpy(&v, b, 1); return v; }
        case PType::I16: { int16_t v; std::memcpy(&v, b, 2); return v; }
        case PType::U16: { uint16_t v; std::memcpy(&v, b, 2); return v; }
        case PType::I32: { int32_t v; std::memcpy(&v, b, 4); return v; }
        case PType::U32: { uint32_t v; std::memcpy(&v, b, 4); return v; }
        case PType::F32: { float v; std::memcpy(&v, b, 4); return v; }
        case PType::F64: { double v; std::memcpy(&v, b, 8); return v; }
    }
    return 0.0;
}

// Position slots: happly's getProperty<double> accepts double and float properties only.
void check_position_type(const Prop& p) {
    if (p.is_list || (p.type != PType::F32 && p.type != PType::F64))
        throw std::runtime_error("PLY parser: property " + p.name +
                                 " cannot be coerced to requested type double");
}
// Colour slots: getProperty<unsigned char> accepts uchar properties only.
void check_color_type(const Prop& p) {
    if (p.is_list || p.type != PType::U8)
        throw std::runtime_error("PLY parser: property " + p.name +
                                 " cannot be coerced to requested type uchar");
}

}  // namespace

PlyCloud load_ply(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("PLY parser: could not open file " + path);
    f.seekg(0, std::ios::end);
    const int64_t file_bytes = (int64_t)f.tellg();
    f.seekg(0, std::ios::beg);
    // bytes left after the current read position (counts in a header are checked against it
    // before anything is allocated for them)
    auto remaining = [&]() -> int64_t {
        const std::streamoff pos = f.tellg();
        return pos < 0 ? 0 : file_bytes - (int64_t)pos;
    };
    std::string line;
    auto next_line = [&](const char* what) {
        if (!std::getline(f, line)) throw std::runtime_error(std::string("PLY parser: ") + what);
        if (!line.empty() && line.back() == '\r') line.pop_back();
    };
    next_line("empty file");
    if (line != "ply") throw std::runtime_error("PLY parser: file " + path + " does not start with 'ply'");
    Format fmt = Format::Ascii;
    bool have_format = false;
    std::vector<Elem> elems;
    for (;;) {
        next_line("header has no end_header");
        std::istringstream ls(line);
        std::string tok;
        ls >> tok;
        if (tok.empty() || tok == "comment" || tok == "obj_info") continue;
        if (tok == "end_header") break;
        if (tok == "format") {
            std::string f_name, ver;
            ls >> f_name >> ver;
            if (f_name == "ascii") fmt = Format::Ascii;
            else if (f_name == "binary_little_endian") fmt = Format::BinLE;
            else if (f_name == "binary_big_endian") fmt = Format::BinBE;
            else throw std::runtime_error("PLY parser: unknown format " + f_name);
            have_format = true;
        } else if (tok == "element") {
            Elem e;
            long long c = -1;
            ls >> e.name >> c;
            if (e.name.empty() || c < 0) throw std::runtime_error("PLY parser: bad element line: " + line);
            e.count = c;
            elems.push_back(e);
        } else if (tok == "property") {
            if (elems.empty()) throw std::runtime_error("PLY parser: property before any element");
            Prop p;
            std::string t1;
            ls >> t1;
            if (t1 == "list") {
                std::string ct, it;
                ls >> ct >> it >> p.name;
                p.is_list = true;
                if (!parse_type(ct, p.count_type) || !parse_type(it, p.type))
                    throw std::runtime_error("PLY parser: bad list property line: " + line);
            } else {
                ls >> p.name;
                if (!parse_type(t1, p.type)) throw std::runtime_error("PLY parser: unknown type " + t1);
            }
            if (p.name.empty()) throw std::runtime_error("PLY parser: property without a name");
            elems.back().props.push_back(p);
        } else {
            throw std::runtime_error("PLY parser: unexpected header line: " + line);
        }
    }
    if (!have_format) throw std::runtime_error("PLY parser: header has no format line");

    int vi = -1;
    for (size_t e = 0; e < elems.size(); ++e)
        if (elems[e].name == "vertex") { vi = (int)e; break; }
    if (vi < 0) throw std::runtime_error("PLY parser: no element with name: vertex");
    const Elem& V = elems[vi];
    const char* want[6] = {"x", "y", "z", "red", "green", "blue"};
    int slot[6];
    for (int k = 0; k < 6; ++k) {
        slot[k] = -1;
        for (size_t p = 0; p < V.props.size(); ++p)
            if (V.props[p].name == want[k]) slot[k] = (int)p;
        if (slot[k] < 0)
            throw std::runtime_error(std::string("PLY parser: element vertex has no property with name: ") + want[k]);
        if (k < 3) check_position_type(V.props[slot[k]]);
        else check_color_type(V.props[slot[k]]);
    }

    // Each vertex takes at least one byte per property value in binary files and two (a digit
    // and a separator) in ascii ones: a count the rest of the file cannot hold is rejected here,
    // before its arrays are allocated.
    {
        int64_t min_bytes = 0;
        for (const Prop& P : V.props)
            min_bytes += fmt == Format::Ascii ? 2 : type_size(P.is_list ? P.count_type : P.type);
        if (min_bytes > 0 && V.count > remaining() / min_bytes)
            throw std::runtime_error(std::string("PLY parser: unexpected end of ") +
                                     (fmt == Format::Ascii ? "ascii" : "binary") +
                                     " data (element vertex count " +
                                     std::to_string(V.count) + " exceeds the file)");
    }
    PlyCloud out;
    out.n = V.count;
    out.xyzw.assign(4 * (size_t)V.count, 0.f);
    out.rgba.assign(4 * (size_t)V.count, 0.f);
    double vals[6];
    auto store = [&](int64_t i) {
        // glm::vec4(double, double, double, 1.0f): each coordinate rounded to float
        out.xyzw[4 * i + 0] = (float)vals[0];
        out.xyzw[4 * i + 1] = (float)vals[1];
        out.xyzw[4 * i + 2] = (float)vals[2];
        out.xyzw[4 * i + 3] = 1.0f;
        out.rgba[4 * i + 0] = (unsigned char)vals[3] / 255.0f;
        out.rgba[4 * i + 1] = (unsigned char)vals[4] / 255.0f;
        out.rgba[4 * i + 2] = (unsigned char)vals[5] / 255.0f;
        out.rgba[4 * i + 3] = 1.0f;
    };

    if (fmt == Format::Ascii) {
        for (size_t e = 0; e < elems.size(); ++e) {
            const Elem& E = elems[e];
            for (int64_t i = 0; i < E.count; ++i) {
                next_line("unexpected end of ascii data");
                if ((int)e != vi) continue;
                std::istringstream ls(line);
                for (size_t p = 0; p < E.props.size(); ++p) {
                    const Prop& P = E.props[p];
                    std::string tok;
                    if (P.is_list) {
                        long long c = 0;
                        if (!(ls >> c) || c < 0) throw std::runtime_error("PLY parser: bad list count");
                        for (long long k = 0; k < c; ++k) ls >> tok;
                        continue;
                    }
                    if (!(ls >> tok)) throw std::runtime_error("PLY parser: missing value in vertex " + std::to_string(i));
                    double v;
                    char* end = nullptr;
                    if (P.type == PType::F32) v = std::strtof(tok.c_str(), &end);   // read as float
                    else if (P.type == PType::F64) v = std::strtod(tok.c_str(), &end);
                    else v = (double)std::strtoll(tok.c_str(), &end, 10);           // ints via long
                    if (end == tok.c_str()) throw std::runtime_error("PLY parser: bad value '" + tok + "'");
                    if (P.type == PType::U8) v = (double)(unsigned char)(long long)v;
                    for (int k = 0; k < 6; ++k)
                        if ((int)p == slot[k]) vals[k] = v;
                }
                store(i);
            }
        }
        return out;
    }

    const bool swap = (fmt == Format::BinBE);
    std::vector<unsigned char> buf;
    auto need = [&](size_t bytes) {
        if ((int64_t)bytes > remaining())
            throw std::runtime_error("PLY parser: unexpected end of binary data");
        buf.resize(bytes);
        f.read((char*)buf.data(), (std::streamsize)bytes);
        if ((size_t)f.gcount() != bytes) throw std::runtime_error("PLY parser: unexpected end of binary data");
    };
    for (size_t e = 0; e < elems.size(); ++e) {
        const Elem& E = elems[e];
        bool fixed = true;
        size_t stride = 0;
        std::vector<size_t> off(E.props.size());
        for (size_t p = 0; p < E.props.size(); ++p) {
            if (E.props[p].is_list) fixed = false;
            off[p] = stride;
            stride += (size_t)type_size(E.props[p].type);
        }
        if (fixed) {
            const int64_t chunk = 1 << 20;
            for (int64_t b = 0; b < E.count; b += chunk) {
                const int64_t cnt = std::min(chunk, E.count - b);
                need((size_t)cnt * stride);
                if ((int)e != vi) continue;
                for (int64_t i = 0; i < cnt; ++i) {
                    const unsigned char* rec = buf.data() + (size_t)i * stride;
                    for (int k = 0; k < 6; ++k)
                        vals[k] = read_scalar(rec + off[slot[k]], E.props[slot[k]].type, swap);
                    store(b + i);
                }
            }
            continue;
        }
        for (int64_t i = 0; i < E.count; ++i) {
            for (size_t p = 0; p < E.props.size(); ++p) {
                const Prop& P = E.props[p];
                if (P.is_list) {
                    need((size_t)type_size(P.count_type));
                    const double c = read_scalar(buf.data(), P.count_type, swap);
                    if (!(c >= 0)) throw std::runtime_error("PLY parser: negative list count");
                    if (c * type_size(P.type) > (double)remaining())
                        throw std::runtime_error("PLY parser: unexpected end of binary data");
                    need((size_t)c * (size_t)type_size(P.type));
                    continue;
                }
                need((size_t)type_size(P.type));
                const double v = read_scalar(buf.data(), P.type, swap);
                if ((int)e == vi)
                    for (int k = 0; k < 6; ++k)
                        if ((int)p == slot[k]) vals[k] = v;
            }
            if ((int)e == vi) store(i);
        }
    }
    return out;
}

void write_ply(const std::string& path, const float* xyz, const uint8_t* rgb, int64_t n) {
    if (n < 0) throw std::runtime_error("write_ply: negative point count");
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("write_ply: could not open " + path);
    f << "ply\nformat binary_little_endian 1.0\nelement vertex " << n
      << "\nproperty float x\nproperty float y\nproperty float z\n"
         "property uchar red\nproperty uchar green\nproperty uchar blue\nend_header\n";
    std::vector<unsigned char> rec((size_t)std::min<int64_t>(n, 1 << 20) * 15);
    for (int64_t b = 0; b < n; b += (1 << 20)) {
        const int64_t cnt = std::min<int64_t>(1 << 20, n - b);
        for (int64_t i = 0; i < cnt; ++i) {
            unsigned char* r = rec.data() + (size_t)i * 15;
            std::memcpy(r, xyz + 3 * (b + i), 12);
            std::memcpy(r + 12, rgb + 3 * (b + i), 3);
        }
        f.write((const char*)rec.data(), (std::streamsize)(cnt * 15));
    }
    if (!f) throw std::runtime_error("write_ply: write failed for " + path);
}

}  // namespace rmh
