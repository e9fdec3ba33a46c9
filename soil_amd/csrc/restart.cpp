// Restart and dump files of soil_print, host side.
//
// RestartData (crates/soil_print/src/lib.rs:317-398) is written with bincode 1
// (`bincode::serialize_into`, :1441) or serde_json (:1436) and read back by
// read_restart (:1527-1610). bincode 1's default encoding, which those calls
// use, is: integers little-endian at their full width (usize as u64), f64 as
// its 8 IEEE bytes, a Vec as a u64 element count followed by the elements, a
// struct as its fields in declaration order with no names or padding.
// serde_json writes the struct as an object keyed by field name.
//
// Dump frames (:1272-1380) are the CSV text, little-endian binary and LAMMPS
// text writers over tag, type, pos, vel, force, radius (+ IsGhost). Rust's
// `{}` prints an f64 as the shortest decimal that reads back to the same value,
// in positional notation, which std::to_chars(..., chars_format::fixed) gives.
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "box.hpp"
#include "soil_hip.h"

namespace soil_hip {

struct RestartHost {
    uint64_t natoms = 0, total_cycle = 0;
    double dt = 0.0;
    std::vector<uint32_t> tag, type;
    std::vector<double> px, py, pz, vx, vy, vz, fx, fy, fz, mass, radius;
    std::vector<std::vector<double>> buffers;
};

// ── bincode 1 ─────────────────────────────────────────────────────────────
struct BinWriter {
    std::vector<unsigned char> out;
    template <typename T>
    void raw(T v) {
        unsigned char b[sizeof(T)];
        std::memcpy(b, &v, sizeof(T));  // little-endian host
        out.insert(out.end(), b, b + sizeof(T));
    }
    template <typename T>
    void vec(const std::vector<T>& v) {
        raw<uint64_t>(v.size());
        for (const T& x : v) raw<T>(x);
    }
};

struct BinReader {
    const unsigned char* p;
    size_t n, off = 0;
    bool ok = true;
    template <typename T>
    T raw() {
        T v{};
        if (off + sizeof(T) > n) {
            ok = false;
            return v;
        }
        std::memcpy(&v, p + off, sizeof(T));
        off += sizeof(T);
        return v;
    }
    template <typename T>
    std::vector<T> vec() {
        const uint64_t len = raw<uint64_t>();
        std::vector<T> v;
        if (!ok || len > (n - off) / sizeof(T)) {
            ok = false;
            return v;
        }
        v.resize(len);
        if (len) std::memcpy(v.data(), p + off, len * sizeof(T));
        off += len * sizeof(T);
        return v;
    }
    void skip_vec(size_t elem) {
        const uint64_t len = raw<uint64_t>();
        if (!ok || len > (n - off) / elem) {
            ok = false;
            return;
        }
        off += len * elem;
    }
};

static std::vector<unsigned char> encode_bincode(const RestartHost& r) {
    BinWriter w;
    w.raw<uint64_t>(r.natoms);
    w.raw<uint64_t>(r.total_cycle);
    w.raw<double>(r.dt);
    w.vec(r.tag);
    w.vec(r.type);
    for (const auto* v : {&r.px, &r.py, &r.pz, &r.vx, &r.vy, &r.vz, &r.fx, &r.fy, &r.fz, &r.mass, &r.radius})
        w.vec(*v);
    w.raw<uint64_t>(r.buffers.size());
    for (const auto& b : r.buffers) w.vec(b);
    for (int k = 0; k < 9; ++k) w.raw<uint64_t>(0);  // omega, torque, ang_mom (x, y, z): empty
    w.raw<uint64_t>(0);                               // quaternion: empty
    return w.out;
}

static bool decode_bincode(const std::vector<unsigned char>& bytes, RestartHost& r) {
    BinReader rd{bytes.data(), bytes.size()};
    r.natoms = rd.raw<uint64_t>();
    r.total_cycle = rd.raw<uint64_t>();
    r.dt = rd.raw<double>();
    r.tag = rd.vec<uint32_t>();
    r.type = rd.vec<uint32_t>();
    for (auto* v : {&r.px, &r.py, &r.pz, &r.vx, &r.vy, &r.vz, &r.fx, &r.fy, &r.fz, &r.mass, &r.radius})
        *v = rd.vec<double>();
    const uint64_t nb = rd.raw<uint64_t>();
    if (!rd.ok || nb > bytes.size() / 8) return false;
    r.buffers.resize(nb);
    for (auto& b : r.buffers) b = rd.vec<double>();
    for (int k = 0; k < 9; ++k) rd.skip_vec(8);  // legacy rotational fields
    rd.skip_vec(32);                              // quaternion [f64; 4]
    return rd.ok;
}

// ── serde_json ───────────────────────────────────────────────────────────
static void put_num(std::string& s, double v) {
    if (!std::isfinite(v)) {  // serde_json writes non-finite floats as null
        s += "null";
        return;
    }
    char b[512];
    auto res = std::to_chars(b, b + sizeof(b), v);  // shortest round-trip form
    s.append(b, res.ptr);
    // a float token, as serde_json writes it (1.0, -0.0): a bare integer would read back as an
    // integer and lose the sign of zero
    if (std::memchr(b, '.', res.ptr - b) == nullptr && std::memchr(b, 'e', res.ptr - b) == nullptr) s += ".0";
}

static void put_key(std::string& s, const char* k) {
    if (s.size() > 1) s += ',';
    s += '"';
    s += k;
    s += "\":";
}

template <typename T>
static void put_array(std::string& s, const std::vector<T>& v) {
    s += '[';
    for (size_t i = 0; i < v.size(); ++i) {
        if (i) s += ',';
        if constexpr (std::is_same<T, double>::value) put_num(s, v[i]);
        else s += std::to_string(v[i]);
    }
    s += ']';
}

static std::string encode_json(const RestartHost& r) {
    std::string s = "{";
    put_key(s, "natoms");
    s += std::to_string(r.natoms);
    put_key(s, "total_cycle");
    s += std::to_string(r.total_cycle);
    put_key(s, "dt");
    put_num(s, r.dt);
    put_key(s, "tag");
    put_array(s, r.tag);
    put_key(s, "atom_type");
    put_array(s, r.type);
    const char* names[11] = {"pos_x", "pos_y", "pos_z", "vel_x", "vel_y", "vel_z",
                             "force_x", "force_y", "force_z", "mass", "cutoff_radius"};
    const std::vector<double>* cols[11] = {&r.px, &r.py, &r.pz, &r.vx, &r.vy, &r.vz,
                                           &r.fx, &r.fy, &r.fz, &r.mass, &r.radius};
    for (int k = 0; k < 11; ++k) {
        put_key(s, names[k]);
        put_array(s, *cols[k]);
    }
    put_key(s, "atom_data_buffers");
    s += '[';
    for (size_t k = 0; k < r.buffers.size(); ++k) {
        if (k) s += ',';
        put_array(s, r.buffers[k]);
    }
    s += ']';
    const char* legacy[10] = {"omega_x", "omega_y", "omega_z", "torque_x", "torque_y",
                              "torque_z", "ang_mom_x", "ang_mom_y", "ang_mom_z", "quaternion"};
    for (const char* k : legacy) {
        put_key(s, k);
        s += "[]";
    }
    s += '}';
    return s;
}

// Minimal JSON reader for this object: numbers, arrays (nested), strings as keys.
struct JsonReader {
    const char* p;
    const char* e;
    bool ok = true;
    void ws() {
        while (p < e && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t')) ++p;
    }
    bool eat(char c) {
        ws();
        if (p < e && *p == c) {
            ++p;
            return true;
        }
        return false;
    }
    bool key(std::string& k) {
        ws();
        if (p >= e || *p != '"') return false;
        ++p;
        const char* s = p;
        while (p < e && *p != '"') {
            if (*p == '\\') return false;  // the writer never escapes field names
            ++p;
        }
        if (p >= e) return false;
        k.assign(s, p);
        ++p;
        return eat(':');
    }
    double number() {
        ws();
        char buf[512];
        size_t n = 0;
        while (p < e && n + 1 < sizeof(buf) && (std::strchr("+-.0123456789eE", *p) != nullptr)) buf[n++] = *p++;
        buf[n] = 0;
        if (n == 0) {
            ok = false;
            return 0.0;
        }
        char* end = nullptr;
        const double v = std::strtod(buf, &end);  // correctly rounded
        if (end != buf + n) ok = false;
        return v;
    }
    uint64_t integer() {
        ws();
        uint64_t v = 0;
        const char* s = p;
        while (p < e && *p >= '0' && *p <= '9') v = v * 10 + (uint64_t)(*p++ - '0');
        if (p == s) ok = false;
        return v;
    }
    template <typename T>
    std::vector<T> array() {
        std::vector<T> v;
        if (!eat('[')) {
            ok = false;
            return v;
        }
        if (eat(']')) return v;
        do {
            if constexpr (std::is_same<T, double>::value) v.push_back(number());
            else v.push_back((T)integer());
        } while (ok && eat(','));
        if (!eat(']')) ok = false;
        return v;
    }
    void skip_value() {  // any array of numbers / arrays, or a number
        ws();
        if (eat('[')) {
            if (eat(']')) return;
            do skip_value();
            while (ok && eat(','));
            if (!eat(']')) ok = false;
        } else {
            number();
        }
    }
};

static bool decode_json(const std::string& text, RestartHost& r, std::string& err) {
    JsonReader j{text.data(), text.data() + text.size()};
    if (!j.eat('{')) {
        err = "not a JSON object";
        return false;
    }
    bool have[17] = {};
    const char* req[17] = {"natoms", "total_cycle", "dt", "tag", "atom_type", "pos_x", "pos_y", "pos_z", "vel_x",
                           "vel_y", "vel_z", "force_x", "force_y", "force_z", "mass", "cutoff_radius",
                           "atom_data_buffers"};
    std::vector<double>* cols[11] = {&r.px, &r.py, &r.pz, &r.vx, &r.vy, &r.vz, &r.fx, &r.fy, &r.fz, &r.mass, &r.radius};
    if (!j.eat('}')) {
        do {
            std::string k;
            if (!j.key(k)) {
                err = "bad key";
                return false;
            }
            int idx = -1;
            for (int q = 0; q < 17; ++q)
                if (k == req[q]) idx = q;
            if (idx == 0) r.natoms = j.integer();
            else if (idx == 1) r.total_cycle = j.integer();
            else if (idx == 2) r.dt = j.number();
            else if (idx == 3) r.tag = j.array<uint32_t>();
            else if (idx == 4) r.type = j.array<uint32_t>();
            else if (idx >= 5 && idx <= 15) *cols[idx - 5] = j.array<double>();
            else if (idx == 16) {
                r.buffers.clear();
                if (!j.eat('[')) {
                    err = "atom_data_buffers is not an array";
                    return false;
                }
                if (!j.eat(']')) {
                    do r.buffers.push_back(j.array<double>());
                    while (j.ok && j.eat(','));
                    if (!j.eat(']')) j.ok = false;
                }
            } else {
                j.skip_value();  // legacy fields and anything else serde would ignore
            }
            if (idx >= 0) have[idx] = true;
            if (!j.ok) {
                err = "bad value for " + k;
                return false;
            }
        } while (j.eat(','));
        if (!j.eat('}')) {
            err = "unterminated object";
            return false;
        }
    }
    for (int q = 0; q < 17; ++q)
        if (!have[q]) {
            err = std::string("missing field ") + req[q];
            return false;
        }
    return true;
}

static bool read_file(const char* path, std::vector<unsigned char>& out) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    unsigned char buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) out.insert(out.end(), buf, buf + n);
    std::fclose(f);
    return true;
}

static bool write_file(const char* path, const void* data, size_t n) {
    FILE* f = std::fopen(path, "wb");
    if (!f) return false;
    const bool ok = std::fwrite(data, 1, n, f) == n;
    return std::fclose(f) == 0 && ok;
}

static int check_restart(const RestartHost& r) {
    const size_t n = r.tag.size();
    const std::vector<double>* cols[11] = {&r.px, &r.py, &r.pz, &r.vx, &r.vy, &r.vz, &r.fx, &r.fy, &r.fz, &r.mass, &r.radius};
    if (r.type.size() != n) return set_error(SOIL_HIP_EINVAL, "restart: atom_type length differs from tag");
    for (auto* c : cols)
        if (c->size() != n) return set_error(SOIL_HIP_EINVAL, "restart: a per-atom vector's length differs from tag");
    return 0;
}

int restart_write(const char* path, int format, const RestartHost& r) {
    int rc = check_restart(r);
    if (rc) return rc;
    if (format == SOIL_HIP_RESTART_JSON) {
        const std::string s = encode_json(r);
        if (!write_file(path, s.data(), s.size())) return set_error(SOIL_HIP_EINVAL, "cannot write %s", path);
    } else if (format == SOIL_HIP_RESTART_BINCODE) {
        const auto b = encode_bincode(r);
        if (!write_file(path, b.data(), b.size())) return set_error(SOIL_HIP_EINVAL, "cannot write %s", path);
    } else {
        return set_error(SOIL_HIP_EINVAL, "unknown restart format %d", format);
    }
    return 0;
}

int restart_read(const char* path, int format, RestartHost& r) {
    std::vector<unsigned char> bytes;
    if (!read_file(path, bytes)) return set_error(SOIL_HIP_EINVAL, "cannot read %s", path);
    if (format == SOIL_HIP_RESTART_JSON) {
        std::string err;
        if (!decode_json(std::string(bytes.begin(), bytes.end()), r, err))
            return set_error(SOIL_HIP_EINVAL, "restart JSON %s: %s", path, err.c_str());
    } else if (format == SOIL_HIP_RESTART_BINCODE) {
        if (!decode_bincode(bytes, r)) return set_error(SOIL_HIP_EINVAL, "restart bincode %s: truncated or malformed", path);
    } else {
        return set_error(SOIL_HIP_EINVAL, "unknown restart format %d", format);
    }
    return check_restart(r);
}

static void from_c(const soil_hip_restart* c, RestartHost& r) {
    const size_t n = c->n;
    r.natoms = c->natoms;
    r.total_cycle = c->total_cycle;
    r.dt = c->dt;
    r.tag.assign(c->tag, c->tag + n);
    r.type.assign(c->type, c->type + n);
    std::vector<double>* xyz[3][3] = {{&r.px, &r.py, &r.pz}, {&r.vx, &r.vy, &r.vz}, {&r.fx, &r.fy, &r.fz}};
    const double* src[3] = {c->pos, c->vel, c->force};
    for (int a = 0; a < 3; ++a)
        for (int d = 0; d < 3; ++d) {
            xyz[a][d]->resize(n);
            for (size_t i = 0; i < n; ++i) (*xyz[a][d])[i] = src[a] ? src[a][3 * i + d] : 0.0;
        }
    r.mass.assign(c->mass, c->mass + n);
    r.radius.assign(c->radius, c->radius + n);
    r.buffers.resize(c->nbuf);
    for (size_t k = 0; k < c->nbuf; ++k) r.buffers[k].assign(c->buf[k], c->buf[k] + c->buf_len[k]);
}

// ── dump frames ───────────────────────────────────────────────────────────
static void put_rust(std::string& s, double v) {
    if (std::isnan(v)) {
        s += "NaN";
        return;
    }
    if (std::isinf(v)) {
        s += v < 0 ? "-inf" : "inf";
        return;
    }
    char b[512];
    auto res = std::to_chars(b, b + sizeof(b), v, std::chars_format::fixed);
    s.append(b, res.ptr);
}

struct Frame {
    uint64_t step = 0;
    size_t n = 0, nlocal = 0;
    bool ghost = false;
    std::vector<uint32_t> tag, type;
    std::vector<double> pos, vel, force, radius;  // AoS-of-3 for pos/vel/force
    double low[3], high[3];
    bool periodic[3];
};

static int write_dump(const std::string& stem, int format, const Frame& f) {
    if (format == SOIL_HIP_DUMP_BINARY) {
        BinWriter w;
        w.raw<uint32_t>((uint32_t)f.n);
        for (size_t i = 0; i < f.n; ++i) {
            w.raw<uint32_t>(f.tag[i]);
            w.raw<uint32_t>(f.type[i]);
            for (int c = 0; c < 3; ++c) w.raw<double>(f.pos[3 * i + c]);
            for (int c = 0; c < 3; ++c) w.raw<double>(f.vel[3 * i + c]);
            for (int c = 0; c < 3; ++c) w.raw<double>(f.force[3 * i + c]);
            w.raw<double>(f.radius[i]);
            if (f.ghost) w.raw<uint8_t>(i >= f.nlocal ? 1 : 0);
        }
        const std::string path = stem + ".bin";
        if (!write_file(path.c_str(), w.out.data(), w.out.size()))
            return set_error(SOIL_HIP_EINVAL, "cannot write %s", path.c_str());
        return 0;
    }
    const bool csv = format == SOIL_HIP_DUMP_TEXT;
    if (!csv && format != SOIL_HIP_DUMP_LAMMPS) return set_error(SOIL_HIP_EINVAL, "unknown dump format %d", format);
    const char sep = csv ? ',' : ' ';
    std::string s;
    if (csv) {
        s += "tag,type,x,y,z,vx,vy,vz,fx,fy,fz,radius";
        if (f.ghost) s += ",IsGhost";
        s += '\n';
    } else {
        s += "ITEM: TIMESTEP\n" + std::to_string(f.step) + "\nITEM: NUMBER OF ATOMS\n" + std::to_string(f.n) + "\n";
        s += "ITEM: BOX BOUNDS";
        for (int d = 0; d < 3; ++d) s += f.periodic[d] ? " pp" : " ff";
        s += '\n';
        for (int d = 0; d < 3; ++d) {
            put_rust(s, f.low[d]);
            s += ' ';
            put_rust(s, f.high[d]);
            s += '\n';
        }
        s += "ITEM: ATOMS id type x y z vx vy vz fx fy fz radius";
        if (f.ghost) s += " IsGhost";
        s += '\n';
    }
    for (size_t i = 0; i < f.n; ++i) {
        s += std::to_string(f.tag[i]);
        s += sep;
        s += std::to_string(f.type[i]);
        for (const std::vector<double>* v : {&f.pos, &f.vel, &f.force})
            for (int c = 0; c < 3; ++c) {
                s += sep;
                put_rust(s, (*v)[3 * i + c]);
            }
        s += sep;
        put_rust(s, f.radius[i]);
        if (f.ghost) {
            s += sep;
            s += i >= f.nlocal ? '1' : '0';
        }
        s += '\n';
    }
    const std::string path = stem + (csv ? ".csv" : ".lammpstrj");
    if (!write_file(path.c_str(), s.data(), s.size())) return set_error(SOIL_HIP_EINVAL, "cannot write %s", path.c_str());
    return 0;
}

}  // namespace soil_hip

using namespace soil_hip;

extern "C" {

int soil_hip_restart_write(const char* path, int format, const soil_hip_restart* r) {
    if (!path || !r) return set_error(SOIL_HIP_EINVAL, "null argument");
    if (r->n && (!r->tag || !r->type || !r->pos || !r->mass || !r->radius))
        return set_error(SOIL_HIP_EINVAL, "restart: tag, type, pos, mass and radius are required");
    if (r->nbuf && (!r->buf || !r->buf_len)) return set_error(SOIL_HIP_EINVAL, "restart: null buffer table");
    RestartHost h;
    from_c(r, h);
    return restart_write(path, format, h);
}

int soil_hip_restart_read(const char* path, int format, soil_hip_restart* r) {
    if (!path || !r) return set_error(SOIL_HIP_EINVAL, "null argument");
    RestartHost h;
    int rc = restart_read(path, format, h);
    if (rc) return rc;
    const size_t n = h.tag.size();
    const size_t cap_n = r->n, cap_buf = r->nbuf;
    r->natoms = h.natoms;
    r->total_cycle = h.total_cycle;
    r->dt = h.dt;
    r->n = n;
    r->nbuf = h.buffers.size();
    if (r->buf_len && cap_buf >= h.buffers.size())
        for (size_t k = 0; k < h.buffers.size(); ++k) r->buf_len[k] = h.buffers[k].size();
    if (!r->tag) return 0;  // size query
    if (cap_n < n || cap_buf < h.buffers.size())
        return set_error(SOIL_HIP_ECAPACITY, "restart holds %zu atoms and %zu buffers", n, h.buffers.size());
    std::memcpy(r->tag, h.tag.data(), n * 4);
    if (r->type) std::memcpy(r->type, h.type.data(), n * 4);
    const std::vector<double>* xyz[3][3] = {{&h.px, &h.py, &h.pz}, {&h.vx, &h.vy, &h.vz}, {&h.fx, &h.fy, &h.fz}};
    double* dst[3] = {r->pos, r->vel, r->force};
    for (int a = 0; a < 3; ++a)
        if (dst[a])
            for (int d = 0; d < 3; ++d)
                for (size_t i = 0; i < n; ++i) dst[a][3 * i + d] = (*xyz[a][d])[i];
    if (r->mass) std::memcpy(r->mass, h.mass.data(), n * 8);
    if (r->radius) std::memcpy(r->radius, h.radius.data(), n * 8);
    if (r->buf)
        for (size_t k = 0; k < h.buffers.size(); ++k)
            if (r->buf[k]) std::memcpy(r->buf[k], h.buffers[k].data(), h.buffers[k].size() * 8);
    return 0;
}

int soil_hip_box_write_restart(soil_hip_box* b, const char* path, int format, uint64_t step) {
    if (!b || !path) return set_error(SOIL_HIP_EINVAL, "null argument");
    uint64_t counts[8];
    int rc = soil_hip_get_counts(b, counts);
    if (rc) return rc;
    const size_t n = counts[0], nall = counts[0] + counts[1];
    std::vector<double> pos(3 * nall), vel(3 * nall), force(3 * nall), rad(nall), mass(nall), inv(nall);
    std::vector<uint32_t> tag(nall), type(nall);
    std::vector<int32_t> origin(nall), image(3 * nall);
    rc = soil_hip_get_atoms(b, nall, pos.data(), vel.data(), force.data(), rad.data(), mass.data(), inv.data(),
                            tag.data(), origin.data(), image.data());
    if (rc) return rc;
    rc = soil_hip_get_types(b, type.data(), nall);
    if (rc) return rc;
    RestartHost h;
    h.natoms = n;  // one rank holds every atom
    h.total_cycle = step;
    h.dt = b->model.dt;
    h.tag.assign(tag.begin(), tag.begin() + n);
    h.type.assign(type.begin(), type.begin() + n);
    std::vector<double>* xyz[3][3] = {{&h.px, &h.py, &h.pz}, {&h.vx, &h.vy, &h.vz}, {&h.fx, &h.fy, &h.fz}};
    const std::vector<double>* src[3] = {&pos, &vel, &force};
    for (int a = 0; a < 3; ++a)
        for (int d = 0; d < 3; ++d) {
            xyz[a][d]->resize(n);
            for (size_t i = 0; i < n; ++i) (*xyz[a][d])[i] = (*src[a])[3 * i + d];
        }
    h.mass.assign(mass.begin(), mass.begin() + n);
    h.radius.assign(rad.begin(), rad.begin() + n);
    // AtomData extensions in registration order, each packed atom by atom (atom.rs:353-365)
    for (size_t c = 0; c < b->cols.size(); ++c) {
        const int w = b->cols[c].width;
        std::vector<double> all((size_t)w * nall);
        rc = soil_hip_get_column(b, (int)c, all.data(), nall);
        if (rc) return rc;
        h.buffers.emplace_back(all.begin(), all.begin() + (size_t)w * n);
    }
    return restart_write(path, format, h);
}

int soil_hip_box_read_restart(soil_hip_box* b, const char* path, int format, uint64_t* step) {
    if (!b || !path) return set_error(SOIL_HIP_EINVAL, "null argument");
    if (b->setup_done || b->nlocal) return set_error(SOIL_HIP_ESTATE, "read a restart into a new box, before upload");
    RestartHost h;
    int rc = restart_read(path, format, h);
    if (rc) return rc;
    const size_t n = h.tag.size();
    if (h.buffers.size() != b->cols.size())
        return set_error(SOIL_HIP_EINVAL, "restart has %zu AtomData buffers, the box registers %zu columns",
                         h.buffers.size(), b->cols.size());
    for (size_t c = 0; c < b->cols.size(); ++c)
        if (h.buffers[c].size() != n * (size_t)b->cols[c].width)
            return set_error(SOIL_HIP_EINVAL, "restart buffer %zu holds %zu values, column width %d needs %zu", c,
                             h.buffers[c].size(), b->cols[c].width, n * (size_t)b->cols[c].width);
    std::vector<double> pos(3 * n), vel(3 * n), force(3 * n);
    for (size_t i = 0; i < n; ++i) {
        pos[3 * i] = h.px[i];
        pos[3 * i + 1] = h.py[i];
        pos[3 * i + 2] = h.pz[i];
        vel[3 * i] = h.vx[i];
        vel[3 * i + 1] = h.vy[i];
        vel[3 * i + 2] = h.vz[i];
        force[3 * i] = h.fx[i];
        force[3 * i + 1] = h.fy[i];
        force[3 * i + 2] = h.fz[i];
    }
    // read_restart (lib.rs:1598-1620): image flags are not stored (zero), dt comes from the file
    rc = soil_hip_upload(b, n, pos.data(), vel.data(), force.data(), h.mass.data(), h.radius.data(), h.tag.data(),
                         h.type.data(), nullptr);
    if (rc) return rc;
    for (size_t c = 0; c < b->cols.size(); ++c) {
        rc = soil_hip_upload_column(b, (int)c, h.buffers[c].data(), n);
        if (rc) return rc;
    }
    b->model.dt = h.dt;
    if (step) *step = h.total_cycle;
    return 0;
}

int soil_hip_box_write_dump(soil_hip_box* b, const char* path_stem, int format, uint64_t step, int ghost) {
    if (!b || !path_stem) return set_error(SOIL_HIP_EINVAL, "null argument");
    uint64_t counts[8];
    int rc = soil_hip_get_counts(b, counts);
    if (rc) return rc;
    Frame f;
    f.step = step;
    f.nlocal = counts[0];
    f.ghost = ghost != 0;
    const size_t nall = counts[0] + counts[1];
    f.n = f.ghost ? nall : f.nlocal;
    std::vector<double> pos(3 * nall), vel(3 * nall), force(3 * nall), rad(nall);
    std::vector<uint32_t> tag(nall), type(nall);
    rc = soil_hip_get_atoms(b, nall, pos.data(), vel.data(), force.data(), rad.data(), nullptr, nullptr, tag.data(),
                            nullptr, nullptr);
    if (rc) return rc;
    rc = soil_hip_get_types(b, type.data(), nall);
    if (rc) return rc;
    f.tag.assign(tag.begin(), tag.begin() + f.n);
    f.type.assign(type.begin(), type.begin() + f.n);
    f.pos.assign(pos.begin(), pos.begin() + 3 * f.n);
    f.vel.assign(vel.begin(), vel.begin() + 3 * f.n);
    f.force.assign(force.begin(), force.begin() + 3 * f.n);
    f.radius.assign(rad.begin(), rad.begin() + f.n);
    for (int d = 0; d < 3; ++d) {
        f.low[d] = b->model.lo[d];
        f.high[d] = b->model.hi[d];
        f.periodic[d] = b->model.periodic[d];
    }
    return write_dump(path_stem, format, f);
}

}  // extern "C"
