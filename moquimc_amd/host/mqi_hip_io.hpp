// Output writers in the reference's on-disk formats, so that tools reading
// moqui's outputs read these unchanged:
//
//   Dij as a scipy-style CSR .npz (io::save_to_npz, mqi_io.hpp:249-320, and
//   the zip writer io::save_npz, mqi_sparse_io.hpp:166-400): an uncompressed
//   zip of indices.npy (uint32), indptr.npy (uint32), shape.npy (uint32
//   [rows, cols]), data.npy (float64, value * scale) and format.npy ("csr");
//   rows = spots, columns = voxels.
//   dense scorers as raw float64 (io::save_to_bin, mqi_io.hpp:165-183) and
//   MetaImage .mhd + .raw / .mha (io::save_to_mhd / save_to_mha,
//   mqi_io.hpp:493-591), with the reference's header text.
//
// Differences by design: CSR rows come from mqi_dij_build (columns sorted,
// the build's threshold applied) where the reference writes its hash-table
// order; a member above 4 GiB - 1 bytes is refused (the reference's 32-bit
// zip fields would wrap); errors throw mqi_hip::io_error instead of
// printing and carrying on.
#ifndef MQI_HIP_IO_HPP
#define MQI_HIP_IO_HPP

#include <array>
#include <cctype>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace mqi_hip
{
namespace io
{

struct io_error : std::runtime_error {
    explicit io_error(const std::string& w) : std::runtime_error(w) {}
};

// CRC-32 (ISO-HDLC: reflected polynomial 0xEDB88320, initial and final
// value 0xFFFFFFFF), the checksum of the zip format, byte-wise with a table
inline uint32_t
crc32_update(uint32_t crc, const void* data, size_t n) {
    static const std::array<uint32_t, 256> table = [] {
        std::array<uint32_t, 256> t{};
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t c = i;
            for (int k = 0; k < 8; ++k) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            t[i] = c;
        }
        return t;
    }();
    const uint8_t* p = static_cast<const uint8_t*>(data);
    crc              = ~crc;
    for (size_t i = 0; i < n; ++i) crc = table[(crc ^ p[i]) & 0xFFu] ^ (crc >> 8);
    return ~crc;
}

// .npy version 1.0 header: magic, version, header length, then the dict
// padded with spaces and a newline to a multiple of 64 bytes
inline std::string
npy_header(const std::string& descr, const std::vector<size_t>& shape) {
    std::ostringstream d;
    d << "{'descr': '" << descr << "', 'fortran_order': False, 'shape': (";
    for (size_t i = 0; i < shape.size(); ++i) d << shape[i] << (shape.size() == 1 ? "," : (i + 1 < shape.size() ? ", " : ""));
    d << "), }";
    std::string dict = d.str();
    const size_t base = 10;
    size_t       len  = dict.size() + 1;
    size_t       pad  = (64 - (base + len) % 64) % 64;
    dict.append(pad, ' ');
    dict.push_back('\n');
    std::string h("\x93NUMPY", 6);
    h.push_back('\x01');
    h.push_back('\x00');
    const uint16_t hl = (uint16_t) dict.size();
    h.push_back((char) (hl & 0xFF));
    h.push_back((char) (hl >> 8));
    return h + dict;
}

// an uncompressed zip archive written member by member
class zip_writer
{
public:
    explicit zip_writer(const std::string& path) : path_(path), f_(path, std::ios::binary | std::ios::trunc) {
        if (!f_) throw io_error("cannot write " + path);
    }

    // one .npy member: header + raw little-endian payload
    void
    add_npy(const std::string& name, const std::string& descr, const std::vector<size_t>& shape, const void* data,
            size_t nbytes) {
        const std::string hdr   = npy_header(descr, shape);
        const uint64_t    total = (uint64_t) hdr.size() + nbytes;
        if (total > 0xFFFFFFFEull) throw io_error(name + " in " + path_ + " exceeds the 4 GiB zip member limit");
        uint32_t crc = crc32_update(0, hdr.data(), hdr.size());
        crc          = crc32_update(crc, data, nbytes);
        entry e{name, crc, (uint32_t) total, (uint32_t) offset_};
        std::string lh;
        put32(lh, 0x04034b50u);   // local file header
        put16(lh, 20);            // version needed
        put16(lh, 0);             // flags
        put16(lh, 0);             // stored
        put16(lh, 0);             // time
        put16(lh, 0);             // date
        put32(lh, crc);
        put32(lh, e.size);
        put32(lh, e.size);
        put16(lh, (uint16_t) name.size());
        put16(lh, 0);
        lh += name;
        write(lh.data(), lh.size());
        write(hdr.data(), hdr.size());
        write(data, nbytes);
        entries_.push_back(e);
    }

    void
    close() {
        std::string cd;
        for (const entry& e : entries_) {
            put32(cd, 0x02014b50u);   // central directory header
            put16(cd, 20);            // version made by
            put16(cd, 20);            // version needed
            put16(cd, 0);
            put16(cd, 0);
            put16(cd, 0);
            put16(cd, 0);
            put32(cd, e.crc);
            put32(cd, e.size);
            put32(cd, e.size);
            put16(cd, (uint16_t) e.name.size());
            put16(cd, 0);   // extra
            put16(cd, 0);   // comment
            put16(cd, 0);   // disk
            put16(cd, 0);   // internal attributes
            put32(cd, 0);   // external attributes
            put32(cd, e.offset);
            cd += e.name;
        }
        const uint64_t cd_off = offset_;
        if (cd_off > 0xFFFFFFFEull) throw io_error(path_ + " exceeds the 4 GiB zip limit");
        std::string end;
        put32(end, 0x06054b50u);   // end of central directory
        put16(end, 0);
        put16(end, 0);
        put16(end, (uint16_t) entries_.size());
        put16(end, (uint16_t) entries_.size());
        put32(end, (uint32_t) cd.size());
        put32(end, (uint32_t) cd_off);
        put16(end, 0);
        write(cd.data(), cd.size());
        write(end.data(), end.size());
        f_.close();
        if (!f_) throw io_error("error writing " + path_);
    }

private:
    struct entry {
        std::string name;
        uint32_t    crc, size, offset;
    };
    static void
    put16(std::string& s, uint16_t v) {
        s.push_back((char) (v & 0xFF));
        s.push_back((char) (v >> 8));
    }
    static void
    put32(std::string& s, uint32_t v) {
        for (int k = 0; k < 4; ++k) s.push_back((char) ((v >> (8 * k)) & 0xFF));
    }
    void
    write(const void* p, size_t n) {
        f_.write(static_cast<const char*>(p), (std::streamsize) n);
        if (!f_) throw io_error("error writing " + path_);
        offset_ += n;
    }

    std::string        path_;
    std::ofstream      f_;
    uint64_t           offset_ = 0;
    std::vector<entry> entries_;
};

// Dij CSR (rows = spots, columns = voxels) as the reference's .npz: data =
// value * scale in float64, uint32 indices and indptr, uint32 shape
inline void
save_npz_csr(const std::string& path, uint32_t rows, uint32_t cols, const int64_t* indptr, const int32_t* indices,
             const float* values, double scale) {
    const uint64_t nnz = (uint64_t) indptr[rows];
    if (nnz > std::numeric_limits<uint32_t>::max()) throw io_error("nnz exceeds the uint32 indptr of the format");
    std::vector<uint32_t> ip(rows + 1), ix(nnz);
    std::vector<double>   dv(nnz);
    for (uint32_t r = 0; r <= rows; ++r) ip[r] = (uint32_t) indptr[r];
    for (uint64_t i = 0; i < nnz; ++i) {
        ix[i] = (uint32_t) indices[i];
        dv[i] = (double) values[i] * scale;
    }
    const uint32_t shape[2] = {rows, cols};
    zip_writer     z(path);
    z.add_npy("indices.npy", "<u4", {(size_t) nnz}, ix.data(), ix.size() * 4);
    z.add_npy("indptr.npy", "<u4", {(size_t) rows + 1}, ip.data(), ip.size() * 4);
    z.add_npy("shape.npy", "<u4", {2}, shape, sizeof shape);
    z.add_npy("data.npy", "<f8", {(size_t) nnz}, dv.data(), dv.size() * 8);
    z.add_npy("format.npy", "|S3", {}, "csr", 3);
    z.close();
}

// dense scorer * scale as raw float64 (<dir>/<name>.raw)
inline void
save_to_bin(const double* src, size_t n, double scale, const std::string& dir, const std::string& name) {
    std::vector<double> d(src, src + n);
    for (double& v : d) v *= scale;
    std::ofstream f(dir + "/" + name + ".raw", std::ios::binary | std::ios::trunc);
    f.write(reinterpret_cast<const char*>(d.data()), (std::streamsize) (n * sizeof(double)));
    if (!f) throw io_error("error writing " + dir + "/" + name + ".raw");
}

// the reference's voxel spacing and first-voxel centre: edge[1] - edge[0]
// and edge[0] + d * 0.5, rounded to float (uniform grids, as there)
inline void
grid_origin(const float* e, float& d, float& o) {
    d = e[1];
    d -= e[0];
    o = (float) (e[0] + d * 0.5);
}

// MetaImage: <dir>/<name>.mhd header + <name>.raw (mha = false), or one
// <dir>/<name>.mha with the data after the header (mha = true)
inline void
save_metaimage(const double* src, int nx, int ny, int nz, const float* xe, const float* ye, const float* ze,
               double scale, const std::string& dir, const std::string& name, bool mha) {
    float dx, dy, dz, x0, y0, z0;
    grid_origin(xe, dx, x0);
    grid_origin(ye, dy, y0);
    grid_origin(ze, dz, z0);
    const size_t        n = (size_t) nx * ny * nz;
    std::vector<double> d(src, src + n);
    for (double& v : d) v *= scale;
    std::ostringstream h;
    h << "ObjectType = Image\n";
    h << "NDims = 3\n";
    h << "BinaryData = True\n";
    h << "BinaryDataByteOrderMSB = False\n";
    h << "CompressedData = False\n";
    if (!mha) {
        h << "TransformMatrix 1 0 0 0 1 0 0 0 1\n";
        h << "Offset " << x0 << " " << y0 << " " << z0 << "\n";
        h << "CenterOfRotation 0 0 0\n";
    } else {
        h << "TransformMatrix = 1 0 0 0 1 0 0 0 1\n";
        h << "Origin = " << std::setprecision(9) << x0 << " " << y0 << " " << z0 << "\n";
        h << "CenterOfRotation = 0 0 0\n";
    }
    h << "AnatomicOrientation = RAI\n";
    h << "DimSize = " << nx << " " << ny << " " << nz << "\n";
    h << "ElementType = MET_DOUBLE\n";
    if (mha) h << "HeaderSize = -1\n";
    h << "ElementSpacing = " << dx << " " << dy << " " << dz << "\n";
    const std::string path = dir + "/" + name + (mha ? ".mha" : ".mhd");
    std::ofstream     f(path, std::ios::binary | std::ios::trunc);
    if (mha) {
        h << "ElementDataFile = LOCAL\n";
        const std::string s = h.str();
        f.write(s.data(), (std::streamsize) s.size());
        f.write(reinterpret_cast<const char*>(d.data()), (std::streamsize) (n * sizeof(double)));
    } else {
        h << "ElementDataFile = " << name << ".raw\n";
        const std::string s = h.str();
        f.write(s.data(), (std::streamsize) s.size());
    }
    if (!f) throw io_error("error writing " + path);
    if (!mha) {
        std::ofstream r(dir + "/" + name + ".raw", std::ios::binary | std::ios::trunc);
        r.write(reinterpret_cast<const char*>(d.data()), (std::streamsize) (n * sizeof(double)));
        if (!r) throw io_error("error writing " + dir + "/" + name + ".raw");
    }
}

// A CT in MetaImage form with the data in the file (ElementDataFile =
// LOCAL), as tps_env::read_ct_image reads it (mqi_tps_env.hpp:615-701 and
// the edge construction :469-524): Offset = centre of the first voxel,
// ElementSpacing = (dx, dy, dz), DimSize = (nx, ny, nz), float32 voxels
// converted to int16 HU by truncation; edges x_i = (x0 - dx / 2) + i dx, y
// likewise, z_i = z_{i-1} + dz from z_0 = z0 - dz / 2, with an optional shift
// added to the first edge of each axis (the volume shift knobs).  Header keys
// are matched without regard to case.  Different on purpose: an ElementType
// other than MET_FLOAT is refused (the reference reads float32 whatever the
// header says), as are a missing key and a short file.
struct ct_image {
    int                  nx = 0, ny = 0, nz = 0;
    std::array<float, 3> origin{}, spacing{};
    std::vector<float>   xe, ye, ze;
    std::vector<int16_t> hu;
};

inline std::string
trim(const std::string& s) {
    const size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    const size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}

inline bool
iequal(const std::string& a, const char* b) {
    size_t n = 0;
    for (; b[n]; ++n)
        if (n >= a.size() || std::tolower((unsigned char) a[n]) != std::tolower((unsigned char) b[n])) return false;
    return n == a.size();
}

inline ct_image
read_mha_ct(const std::string& path, const float* shift = nullptr) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw io_error("cannot open " + path);
    ct_image    ct;
    bool        have_off = false, have_sp = false, have_dim = false, local = false;
    std::string line;
    while (std::getline(f, line)) {
        const size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        const std::string key = trim(line.substr(0, eq));
        std::istringstream v(trim(line.substr(eq + 1)));
        // numbers are read as double and rounded to float, as atof into a
        // float does
        double a, b, c;
        if (iequal(key, "Offset")) {
            have_off  = (bool) (v >> a >> b >> c);
            ct.origin = {(float) a, (float) b, (float) c};
        } else if (iequal(key, "ElementSpacing")) {
            have_sp    = (bool) (v >> a >> b >> c);
            ct.spacing = {(float) a, (float) b, (float) c};
        } else if (iequal(key, "DimSize")) {
            have_dim = (bool) (v >> ct.nx >> ct.ny >> ct.nz);
        } else if (iequal(key, "ElementType")) {
            std::string t;
            v >> t;
            if (t != "MET_FLOAT") throw io_error(path + ": ElementType " + t + " (the CT reader takes MET_FLOAT)");
        } else if (iequal(key, "ElementDataFile")) {
            std::string t;
            v >> t;
            if (!iequal(t, "LOCAL")) throw io_error(path + ": the data must follow the header (ElementDataFile = LOCAL)");
            local = true;
            break;
        }
    }
    if (!have_off || !have_sp || !have_dim || !local) throw io_error(path + ": Offset, ElementSpacing, DimSize or data missing");
    if (ct.nx <= 0 || ct.ny <= 0 || ct.nz <= 0) throw io_error(path + ": bad DimSize");
    const size_t       n = (size_t) ct.nx * ct.ny * ct.nz;
    std::vector<float> raw(n);
    f.read(reinterpret_cast<char*>(raw.data()), (std::streamsize) (n * sizeof(float)));
    if ((size_t) f.gcount() != n * sizeof(float)) throw io_error(path + ": fewer voxels than DimSize");
    ct.hu.resize(n);
    for (size_t i = 0; i < n; ++i) ct.hu[i] = (int16_t) raw[i];
    const float dx = ct.spacing[0], dy = ct.spacing[1], dz = ct.spacing[2];
    float       x0 = ct.origin[0] - dx / 2.0;
    float       y0 = ct.origin[1] - dy / 2.0;
    float       z0 = ct.origin[2] - dz / 2.0;
    if (shift) {
        x0 += shift[0];
        y0 += shift[1];
        z0 += shift[2];
    }
    ct.xe.resize(ct.nx + 1);
    ct.ye.resize(ct.ny + 1);
    ct.ze.resize(ct.nz + 1);
    for (int i = 0; i <= ct.nx; ++i) ct.xe[i] = x0 + i * dx;
    for (int i = 0; i <= ct.ny; ++i) ct.ye[i] = y0 + i * dy;
    ct.ze[0] = z0;
    for (int i = 1; i <= ct.nz; ++i) ct.ze[i] = ct.ze[i - 1] + dz;
    return ct;
}

// A structure mask in MetaImage form, as mask_reader::read_mha_file reads
// it (mqi_file_handler.hpp:39-89): DimSize, the data after the header
// (ElementDataFile = LOCAL), one byte per voxel, x fastest.  Different on
// purpose: the mask must have the grid's dimensions (the reference reads
// the mask's own DimSize worth of bytes and then indexes it over the CT),
// an ElementType other than MET_UCHAR / MET_CHAR is refused, and so is a
// value other than 0 or 1 (the reference asserts 0 <= v <= 1).
inline std::vector<uint8_t>
read_mask_mha(const std::string& path, int nx, int ny, int nz) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw io_error("cannot open " + path);
    int         mx = 0, my = 0, mz = 0;
    bool        have_dim = false, local = false;
    std::string line;
    while (std::getline(f, line)) {
        const size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        const std::string  key = trim(line.substr(0, eq));
        std::istringstream v(trim(line.substr(eq + 1)));
        if (iequal(key, "DimSize")) {
            have_dim = (bool) (v >> mx >> my >> mz);
        } else if (iequal(key, "ElementType")) {
            std::string t;
            v >> t;
            if (t != "MET_UCHAR" && t != "MET_CHAR")
                throw io_error(path + ": ElementType " + t + " (a mask is MET_UCHAR)");
        } else if (iequal(key, "ElementDataFile")) {
            std::string t;
            v >> t;
            if (!iequal(t, "LOCAL")) throw io_error(path + ": the data must follow the header (ElementDataFile = LOCAL)");
            local = true;
            break;
        }
    }
    if (!have_dim || !local) throw io_error(path + ": DimSize or data missing");
    if (mx != nx || my != ny || mz != nz) throw io_error(path + ": DimSize differs from the grid's");
    const size_t         n = (size_t) nx * ny * nz;
    std::vector<uint8_t> m(n);
    f.read(reinterpret_cast<char*>(m.data()), (std::streamsize) n);
    if ((size_t) f.gcount() != n) throw io_error(path + ": fewer voxels than DimSize");
    for (size_t i = 0; i < n; ++i)
        if (m[i] > 1) throw io_error(path + ": mask values must be 0 or 1");
    return m;
}

// mask_reader::read_mask_files (mqi_file_handler.hpp:92-113): the sum of
// the structure masks (mask_total), the input of mask_to_roi
inline std::vector<uint8_t>
read_mask_total(const std::vector<std::string>& paths, int nx, int ny, int nz) {
    if (paths.empty()) throw io_error("no mask files");
    std::vector<uint8_t> total((size_t) nx * ny * nz, 0);
    for (const std::string& p : paths) {
        const std::vector<uint8_t> m = read_mask_mha(p, nx, ny, nz);
        for (size_t i = 0; i < total.size(); ++i) total[i] = (uint8_t) (total[i] + m[i]);
    }
    return total;
}

}   // namespace io
}   // namespace mqi_hip

#endif
