// DICOM CT series writer for the final stack (int16 stored values + rescale).
//
// The reference writes the final stack as one raw float32 file
// (src/setup.cu:769-787); a DICOM series with HU stored as int16 is the output
// format that follows it.  Each slice becomes one Part-10 file: a 128-byte
// preamble, "DICM", the file meta group (0002) and the data set, all in the
// explicit VR little endian transfer syntax, elements in ascending tag order.
// Only the attributes a CT image needs to be displayed and placed are written
// (SOP common, patient/study/series identification left empty where DICOM
// allows, image plane, image pixel, CT rescale).
#include <sys/types.h>
#include <unistd.h>

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ctbb_hip.h"

namespace {

const char *kCtImageStorage = "1.2.840.10008.5.1.4.1.1.2";
const char *kExplicitVrLittle = "1.2.840.10008.1.2.1";

class Writer {
  public:
    std::vector<uint8_t> b;

    void u16(uint16_t v) {
        b.push_back(static_cast<uint8_t>(v & 0xff));
        b.push_back(static_cast<uint8_t>(v >> 8));
    }
    void u32(uint32_t v) {
        u16(static_cast<uint16_t>(v & 0xffff));
        u16(static_cast<uint16_t>(v >> 16));
    }
    void head(uint16_t g, uint16_t e, const char *vr, uint32_t len) {
        u16(g);
        u16(e);
        b.push_back(static_cast<uint8_t>(vr[0]));
        b.push_back(static_cast<uint8_t>(vr[1]));
        const bool long_form = !std::strcmp(vr, "OB") || !std::strcmp(vr, "OW") || !std::strcmp(vr, "UN") ||
                               !std::strcmp(vr, "SQ") || !std::strcmp(vr, "UT");
        if (long_form) {
            u16(0);
            u32(len);
        } else {
            u16(static_cast<uint16_t>(len));
        }
    }
    // string value padded to even length (UI with NUL, text with space)
    void str(uint16_t g, uint16_t e, const char *vr, const std::string &v) {
        std::string s = v;
        if (s.size() & 1) s.push_back(std::strcmp(vr, "UI") == 0 ? '\0' : ' ');
        head(g, e, vr, static_cast<uint32_t>(s.size()));
        b.insert(b.end(), s.begin(), s.end());
    }
    void us(uint16_t g, uint16_t e, uint16_t v) {
        head(g, e, "US", 2);
        u16(v);
    }
    void ul(uint16_t g, uint16_t e, uint32_t v) {
        head(g, e, "UL", 4);
        u32(v);
    }
};

// Decimal string of a DS value: at most 16 characters.
std::string ds(double v) {
    char buf[32];
    for (int prec = 10; prec >= 1; prec--) {
        std::snprintf(buf, sizeof(buf), "%.*g", prec, v);
        if (std::strlen(buf) <= 16) return buf;
    }
    return "0";
}

std::string ds3(double a, double b, double c) { return ds(a) + "\\" + ds(b) + "\\" + ds(c); }

// 64-bit mix (splitmix64 finaliser)
uint64_t mix(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

// UID under the 2.25 root (a UUID-sized integer in decimal, ITU-T X.667):
// 128 bits from the clock, the process and a per-call counter.
std::string new_uid(uint64_t salt) {
    static uint64_t counter = 0;
    const uint64_t t = static_cast<uint64_t>(std::chrono::high_resolution_clock::now().time_since_epoch().count());
    const uint64_t hi = mix(t ^ (static_cast<uint64_t>(getpid()) << 32) ^ salt);
    const uint64_t lo = mix(hi ^ (++counter) ^ (salt << 1));
    unsigned __int128 v = (static_cast<unsigned __int128>(hi) << 64) | lo;
    v >>= 1;  // 127 bits: at most 39 digits, no leading zero
    std::string d;
    do {
        d.insert(d.begin(), static_cast<char>('0' + static_cast<int>(v % 10)));
        v /= 10;
    } while (v != 0);
    return "2.25." + d;
}

}  // namespace

extern "C" int ctbb_write_dicom_series(const struct recon_metadata *mr, const int16_t *h_stack, float slope,
                                       float intercept, const char *dir, const char *prefix) {
    if (!mr || !h_stack || !dir || !prefix) return CTBB_ERR_ARG;
    if (!(slope > 0.0f) || !std::isfinite(slope) || !std::isfinite(intercept)) return CTBB_ERR_ARG;
    const recon_params &rp = mr->rp;
    const int nf = ctbb_n_slices_final(mr);
    if (nf <= 0 || rp.nx == 0 || rp.ny == 0 || rp.nx > 65535 || rp.ny > 65535) return CTBB_ERR_SHAPE;
    const size_t nxy = static_cast<size_t>(rp.nx) * rp.ny;
    if (nxy * 2 > 0xfffffffeull) return CTBB_ERR_SHAPE;

    // voxel centres (reference include/backproject.cuh:59-64): x index i along
    // sx * dx, y index j along +dy, origin (x_origin, y_origin)
    const double dx = static_cast<double>(rp.recon_fov) / rp.nx, dy = static_cast<double>(rp.recon_fov) / rp.ny;
    const double sx = mr->cg.table_direction == -1 ? 1.0 : -1.0;
    const double x0 = sx * dx * (0.0 - (rp.nx - 1) / 2.0) + rp.x_origin;
    const double y0 = dy * (0.0 - (rp.ny - 1) / 2.0) + rp.y_origin;
    const float d = rp.end_pos - rp.start_pos;
    const int recon_direction = d < 0.0f ? -1 : 1;

    const std::string study = new_uid(1), series = new_uid(2), frame = new_uid(3);
    const std::string impl = "2.25.329800735698586629295641978511506172918";
    for (int k = 0; k < nf; k++) {
        const std::string inst = series + "." + std::to_string(k + 1);
        const double z = static_cast<double>(rp.start_pos) + recon_direction * k * static_cast<double>(rp.slice_thickness);

        Writer meta;
        meta.head(0x0002, 0x0001, "OB", 2);
        meta.b.push_back(0);
        meta.b.push_back(1);
        meta.str(0x0002, 0x0002, "UI", kCtImageStorage);
        meta.str(0x0002, 0x0003, "UI", inst);
        meta.str(0x0002, 0x0010, "UI", kExplicitVrLittle);
        meta.str(0x0002, 0x0012, "UI", impl);
        meta.str(0x0002, 0x0013, "SH", "CTBB_HIP");

        Writer w;
        w.b.assign(128, 0);
        w.b.insert(w.b.end(), {'D', 'I', 'C', 'M'});
        w.ul(0x0002, 0x0000, static_cast<uint32_t>(meta.b.size()));
        w.b.insert(w.b.end(), meta.b.begin(), meta.b.end());
        w.str(0x0008, 0x0008, "CS", "DERIVED\\SECONDARY\\AXIAL");
        w.str(0x0008, 0x0016, "UI", kCtImageStorage);
        w.str(0x0008, 0x0018, "UI", inst);
        w.str(0x0008, 0x0060, "CS", "CT");
        w.str(0x0008, 0x103e, "LO", std::string(prefix).substr(0, 64));
        w.str(0x0010, 0x0010, "PN", "");
        w.str(0x0010, 0x0020, "LO", "");
        w.str(0x0018, 0x0050, "DS", ds(rp.slice_thickness));
        w.str(0x0020, 0x000d, "UI", study);
        w.str(0x0020, 0x000e, "UI", series);
        w.str(0x0020, 0x0011, "IS", "1");
        w.str(0x0020, 0x0013, "IS", std::to_string(k + 1));
        w.str(0x0020, 0x0032, "DS", ds3(x0, y0, z));
        // row direction: along a row the column index j (y) grows; column
        // direction: down a column the row index i (x) grows
        w.str(0x0020, 0x0037, "DS", ds3(0.0, 1.0, 0.0) + "\\" + ds3(sx, 0.0, 0.0));
        w.str(0x0020, 0x0052, "UI", frame);
        w.str(0x0020, 0x1041, "DS", ds(z));
        w.us(0x0028, 0x0002, 1);
        w.str(0x0028, 0x0004, "CS", "MONOCHROME2");
        w.us(0x0028, 0x0010, static_cast<uint16_t>(rp.nx));
        w.us(0x0028, 0x0011, static_cast<uint16_t>(rp.ny));
        w.str(0x0028, 0x0030, "DS", ds(dx) + "\\" + ds(dy));
        w.us(0x0028, 0x0100, 16);
        w.us(0x0028, 0x0101, 16);
        w.us(0x0028, 0x0102, 15);
        w.us(0x0028, 0x0103, 1);
        w.str(0x0028, 0x1050, "DS", "40");
        w.str(0x0028, 0x1051, "DS", "400");
        w.str(0x0028, 0x1052, "DS", ds(intercept));
        w.str(0x0028, 0x1053, "DS", ds(slope));
        w.str(0x0028, 0x1054, "LO", "HU");
        w.head(0x7fe0, 0x0010, "OW", static_cast<uint32_t>(nxy * 2));
        const int16_t *px = h_stack + static_cast<size_t>(k) * nxy;
        const size_t at = w.b.size();
        w.b.resize(at + nxy * 2);
        for (size_t i = 0; i < nxy; i++) {
            const uint16_t u = static_cast<uint16_t>(px[i]);
            w.b[at + 2 * i] = static_cast<uint8_t>(u & 0xff);
            w.b[at + 2 * i + 1] = static_cast<uint8_t>(u >> 8);
        }

        char name[32];
        std::snprintf(name, sizeof(name), "_%04d.dcm", k + 1);
        const std::string path = std::string(dir) + "/" + prefix + name;
        FILE *f = std::fopen(path.c_str(), "wb");
        if (!f) return CTBB_ERR_IO;
        const bool ok = std::fwrite(w.b.data(), 1, w.b.size(), f) == w.b.size();
        if (std::fclose(f) != 0 || !ok) return CTBB_ERR_IO;
    }
    return CTBB_OK;
}
