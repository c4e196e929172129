// Host-side planning and I/O of the reconstruction: prm parsing, scanner
// geometry, slice/block planning, binary raw reader and output writer.
//
// The arithmetic deliberately follows the reference's own float/double/int
// mix (same operation order and types) so that every derived shape, pull
// index and slice position is identical to the reference program's; each
// function cites the reference lines it restates.  Errors are returned as
// ctbb_status codes instead of exit().
#include <dlfcn.h>
#include <pwd.h>
#include <sys/types.h>
#include <unistd.h>

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "ctbb_hip.h"
#include "ctbb_internal.h"

namespace {

constexpr float kPiTrue = 3.1415926535897f;   // reference src/setup.cu:33
constexpr int kBlockSlices = 32;              // reference src/setup.cu:34

void remove_trailing_slash(char *s) {
    size_t n = std::strlen(s);
    if (n > 0 && s[n - 1] == '/') s[n - 1] = '\0';
}

bool file_exists(const char *p) { return access(p, F_OK) != -1; }

// Directory that holds resources/: CTBB_INSTALL_DIR, else the directory of
// this shared library.
std::string library_dir() {
    const char *env = std::getenv("CTBB_INSTALL_DIR");
    if (env && *env) return std::string(env);
    Dl_info info;
    if (dladdr(reinterpret_cast<void *>(&library_dir), &info) && info.dli_fname) {
        std::string p(info.dli_fname);
        size_t s = p.find_last_of('/');
        if (s != std::string::npos) return p.substr(0, s);
    }
    char buf[4096] = {0};
    if (getcwd(buf, sizeof(buf))) return std::string(buf);
    return std::string(".");
}

int direction_of(double a, double b) {
    // int(fabs(d)/d) as in reference src/setup.cu:427,507 — d==0 is treated
    // as +1 (the reference would divide 0/0).
    double d = b - a;
    if (d == 0.0) return 1;
    return static_cast<int>(std::fabs(d) / d);
}

int recon_direction_of(float start, float end) {
    // reference src/setup.cu:486-488
    float d = end - start;
    if (d == 0.0f) return 1;
    int r = static_cast<int>(std::fabs(d) / d);
    if (r != 1 && r != -1) r = 1;
    return r;
}

// Rounds v up to the next multiple of m using C remainder semantics, exactly
// as "(v-1)+(m-(v-1)%m)" in reference src/setup.cu:564-580.
int round_up_c(int v, int m) { return (v - 1) + (m - (v - 1) % m); }

struct PrmToken {
    const char *key;
    char kind;    // 's' string, 'd' int, 'f' float, 'u' unsigned
    size_t off;
    size_t cap;
};

#define PRM_FIELD(k, kind, field) \
    { k, kind, offsetof(recon_params, field), sizeof(((recon_params *)0)->field) }

// Token table: reference resources/config_dict.txt:4-33 (same order, so the
// chained matching behaves the same).
const PrmToken kPrmTokens[] = {
    PRM_FIELD("RawDataDir:", 's', raw_data_dir),
    PRM_FIELD("RawDataFile:", 's', raw_data_file),
    PRM_FIELD("OutputDir:", 's', output_dir),
    PRM_FIELD("OutputFile:", 's', output_file),
    PRM_FIELD("Nrows:", 'd', n_rows),
    PRM_FIELD("CollSlicewidth:", 'f', coll_slicewidth),
    PRM_FIELD("StartPos:", 'f', start_pos),
    PRM_FIELD("EndPos:", 'f', end_pos),
    PRM_FIELD("TableFeed:", 'f', pitch_value),
    PRM_FIELD("PitchValue:", 'f', pitch_value),
    PRM_FIELD("SliceThickness:", 'f', slice_thickness),
    PRM_FIELD("AcqFOV:", 'f', acq_fov),
    PRM_FIELD("ReconFOV:", 'f', recon_fov),
    PRM_FIELD("ReconKernel:", 'd', recon_kernel),
    PRM_FIELD("Readings:", 'd', n_readings),
    PRM_FIELD("Xorigin:", 'f', x_origin),
    PRM_FIELD("Yorigin:", 'f', y_origin),
    PRM_FIELD("Zffs:", 'd', z_ffs),
    PRM_FIELD("Phiffs:", 'd', phi_ffs),
    PRM_FIELD("Scanner:", 's', scanner),
    PRM_FIELD("FileType:", 'd', file_type),
    PRM_FIELD("FileSubType:", 'd', file_subtype),
    PRM_FIELD("RawOffset:", 'd', raw_data_offset),
    PRM_FIELD("Nx:", 'u', nx),
    PRM_FIELD("Ny:", 'u', ny),
    PRM_FIELD("TubeStartAngle:", 'f', tube_start_angle),
    PRM_FIELD("AdaptiveFiltration:", 'f', adaptive_filtration_s),
    PRM_FIELD("NSlices:", 'd', n_slices),
    PRM_FIELD("TableDir:", 's', table_dir_str),
    PRM_FIELD("TableDirInt:", 'd', table_dir),
};

void scan_value(const PrmToken &t, const char *value, recon_params *p) {
    char *base = reinterpret_cast<char *>(p) + t.off;
    switch (t.kind) {
        case 's': {
            // "%s" reads one whitespace-delimited word
            char word[4352] = {0};
            if (std::sscanf(value, "%4351s", word) == 1) {
                std::strncpy(base, word, t.cap - 1);
                base[t.cap - 1] = '\0';
            }
            break;
        }
        case 'd': std::sscanf(value, "%d", reinterpret_cast<int *>(base)); break;
        case 'u': std::sscanf(value, "%u", reinterpret_cast<unsigned *>(base)); break;
        case 'f': std::sscanf(value, "%f", reinterpret_cast<float *>(base)); break;
    }
}

}  // namespace

extern "C" {

const char *ctbb_strerror(int s) {
    switch (s) {
        case CTBB_OK: return "success";
        case CTBB_ERR_ARG: return "invalid argument";
        case CTBB_ERR_IO: return "file missing, unreadable or unwritable";
        case CTBB_ERR_PRM: return "essential prm value missing";
        case CTBB_ERR_RANGE: return "requested reconstruction outside allowed data range";
        case CTBB_ERR_SCANNER: return "could not parse selected scanner";
        case CTBB_ERR_FILTER: return "filter file not found";
        case CTBB_ERR_SHAPE: return "geometry not supported by the kernels";
        case CTBB_ERR_HIP: return "HIP runtime error";
        case CTBB_ERR_NOMEM: return "allocation failed";
        case CTBB_ERR_UNSUPPORTED: return "raw file type not supported by this build";
        case CTBB_ERR_LIMIT: return "input beyond the fixed capacity of a no-allocation entry point";
        case CTBB_ERR_NONFINITE: return "raw data holds NaN or infinite samples";
    }
    return "unknown status";
}

// reference src/parse_config.cu:40-216
int ctbb_parse_config(const char *config_file, struct recon_params *structure) {
    if (!config_file || !structure) return CTBB_ERR_ARG;
    FILE *conf = std::fopen(config_file, "r");
    if (!conf) return CTBB_ERR_IO;
    char raw_line[1024];
    char line[1024];
    std::memset(line, 0, sizeof(line));
    while (std::fgets(raw_line, sizeof(raw_line), conf) != nullptr) {
        size_t cut = std::strcspn(raw_line, "%");   // strip '%' comments
        std::memcpy(line, raw_line, cut);
        line[cut] = '\0';
        if (std::strcmp(line, "") != 0 && std::strcmp(line, "\n") != 0) {
            char *save = nullptr;
            char *token = strtok_r(line, " \t", &save);
            for (const PrmToken &t : kPrmTokens) {
                if (token && std::strcmp(token, t.key) == 0) {
                    if (std::getenv("CTBB_VERBOSE")) std::printf("Token \"%s\" detected\n", token);
                    token = strtok_r(nullptr, " \t", &save);
                    if (token) scan_value(t, token, structure);
                }
            }
        }
        std::memset(line, 0, sizeof(line));
    }
    std::fclose(conf);
    return CTBB_OK;
}

// reference include/parse_config.h:69-113
int ctbb_empty_config(const char *filepath) {
    FILE *f = std::fopen(filepath, "w");
    if (!f) return CTBB_ERR_IO;
    for (const PrmToken &t : kPrmTokens) std::fprintf(f, "%s\n", t.key);
    std::fclose(f);
    return CTBB_OK;
}

int ctbb_prm_float(const char *prm_file, const char *token, float *value) {
    if (!prm_file || !token || !value) return CTBB_ERR_ARG;
    *value = 0.0f;
    FILE *f = std::fopen(prm_file, "r");
    if (!f) return CTBB_ERR_IO;
    char raw_line[1024];
    while (std::fgets(raw_line, sizeof(raw_line), f)) {
        raw_line[std::strcspn(raw_line, "%")] = '\0';
        char *save = nullptr;
        char *tok = strtok_r(raw_line, " \t\r\n", &save);
        if (tok && std::strcmp(tok, token) == 0) {
            tok = strtok_r(nullptr, " \t\r\n", &save);
            if (tok) std::sscanf(tok, "%f", value);
        }
    }
    std::fclose(f);
    return CTBB_OK;
}

int ctbb_prm_mu_water(const char *prm_file, float *mu_water) {
    return ctbb_prm_float(prm_file, "MuWater:", mu_water);
}

// reference src/setup.cu:108-176
int ctbb_configure_recon_params(const char *prm_file, struct recon_params *out) {
    if (!prm_file || !out) return CTBB_ERR_ARG;
    if (!file_exists(prm_file)) return CTBB_ERR_IO;
    recon_params prms;
    std::memset(&prms, 0, sizeof(prms));
    int st = ctbb_parse_config(prm_file, &prms);
    if (st) return st;
    if (std::strcmp(prms.table_dir_str, "") != 0) {
        if (std::strcmp(prms.table_dir_str, "out") == 0) prms.table_dir = 1;
        else if (std::strcmp(prms.table_dir_str, "in") == 0) prms.table_dir = -1;
        else prms.table_dir = 1;
    }
    int bad = 0;
    bad |= prms.n_rows == 0;
    bad |= prms.coll_slicewidth == 0;
    bad |= prms.slice_thickness == 0;
    bad |= prms.pitch_value == 0;
    bad |= prms.acq_fov == 0;
    bad |= prms.recon_fov == 0;
    bad |= prms.n_readings == 0;
    bad |= prms.nx == 0;
    bad |= prms.ny == 0;
    if (prms.file_type == 0 && prms.table_dir == 0) prms.table_dir = 1;
    *out = prms;
    return bad ? CTBB_ERR_PRM : CTBB_OK;
}

// reference src/setup.cu:40-106
int ctbb_configure_paths(struct recon_metadata *mr) {
    if (!mr) return CTBB_ERR_ARG;
    struct passwd *pw = getpwuid(getuid());
    if (pw && pw->pw_dir) std::snprintf(mr->home_dir, sizeof(mr->home_dir), "%s", pw->pw_dir);
    if (!getcwd(mr->cwd, sizeof(mr->cwd))) mr->cwd[0] = '\0';
    std::snprintf(mr->install_dir, sizeof(mr->install_dir), "%s", library_dir().c_str());
    if (std::strcmp(mr->rp.output_dir, "") == 0)
        std::snprintf(mr->rp.output_dir, sizeof(mr->rp.output_dir), "%s", mr->cwd);
    if (std::strcmp(mr->rp.output_file, "") == 0) {
        char full[4096 + 255] = {0};
        std::snprintf(full, sizeof(full), "%s.img", mr->rp.raw_data_file);
        std::snprintf(mr->rp.output_file, sizeof(mr->rp.output_file), "%s", full);
    }
    remove_trailing_slash(mr->home_dir);
    remove_trailing_slash(mr->install_dir);
    remove_trailing_slash(mr->cwd);
    remove_trailing_slash(mr->rp.output_dir);
    remove_trailing_slash(mr->rp.raw_data_dir);

    char full[4096 + 255 + 2] = {0};
    std::snprintf(full, sizeof(full), "%s/%s", mr->rp.raw_data_dir, mr->rp.raw_data_file);
    FILE *f = std::fopen(full, "r");
    if (!f) return 1;
    std::fclose(f);
    std::snprintf(full, sizeof(full), "%s/%s", mr->rp.output_dir, mr->rp.output_file);
    bool existed = file_exists(full);
    f = std::fopen(full, "a");
    if (!f) return 2;
    std::fclose(f);
    if (!existed) std::remove(full);
    return 0;
}

// reference src/setup.cu:178-383
int ctbb_configure_ct_geom(struct recon_metadata *mr, struct ct_geom *out) {
    if (!mr || !out) return CTBB_ERR_ARG;
    ct_geom cg;
    std::memset(&cg, 0, sizeof(cg));
    const recon_params &rp = mr->rp;
    cg.table_direction = rp.table_dir;
    int scanner = -1;

    FILE *f = std::fopen(mr->rp.scanner, "r");
    if (!f) {
        std::string path = std::string(mr->install_dir) + "/resources/scanners/" + mr->rp.scanner;
        f = std::fopen(path.c_str(), "r");
        if (!f) {
            scanner = std::atoi(mr->rp.scanner);
            if (scanner < 0 || scanner > 2 || (scanner == 0 && std::strlen(mr->rp.scanner) != 1))
                return CTBB_ERR_SCANNER;
            if (scanner == 1) mr->rp.file_subtype = 1;
            if (scanner == 2) mr->rp.file_subtype = 2;
        }
    }

    if (scanner == -1) {
        std::fseek(f, 0, SEEK_END);
        long sz = std::ftell(f);
        std::rewind(f);
        std::string buf(static_cast<size_t>(sz) + 1, '\0');
        size_t got = std::fread(&buf[0], 1, static_cast<size_t>(sz), f);
        buf[got] = '\0';
        std::fclose(f);
        const char *delim = " \t\n%";
        char *save = nullptr;
        char *tok = strtok_r(&buf[0], delim, &save);
        while (tok) {
            auto next = [&]() { return strtok_r(nullptr, delim, &save); };
            if (std::strcmp(tok, "RSrcToIso:") == 0) { tok = next(); if (tok) std::sscanf(tok, "%f", &cg.r_f); }
            else if (std::strcmp(tok, "RSrcToDet:") == 0) { tok = next(); if (tok) std::sscanf(tok, "%f", &cg.src_to_det); }
            else if (std::strcmp(tok, "AnodeAngle:") == 0) { tok = next(); if (tok) std::sscanf(tok, "%f", &cg.anode_angle); }
            else if (std::strcmp(tok, "FanAngleInc:") == 0) { tok = next(); if (tok) std::sscanf(tok, "%f", &cg.fan_angle_increment); }
            else if (std::strcmp(tok, "ThetaCone:") == 0) { tok = next(); if (tok) std::sscanf(tok, "%f", &cg.theta_cone); }
            else if (std::strcmp(tok, "CentralChannel:") == 0) { tok = next(); if (tok) std::sscanf(tok, "%f", &cg.central_channel); }
            else if (std::strcmp(tok, "NProjTurn:") == 0) { tok = next(); if (tok) std::sscanf(tok, "%zu", &cg.n_proj_turn); }
            else if (std::strcmp(tok, "NChannels:") == 0) { tok = next(); if (tok) std::sscanf(tok, "%zu", &cg.n_channels); }
            else if (std::strcmp(tok, "ReverseRowInterleave:") == 0) { tok = next(); if (tok) std::sscanf(tok, "%i", &cg.reverse_row_interleave); }
            else if (std::strcmp(tok, "ReverseChanInterleave:") == 0) { tok = next(); if (tok) std::sscanf(tok, "%i", &cg.reverse_channel_interleave); }
            if (!tok) break;
            tok = strtok_r(nullptr, delim, &save);
        }
    } else {
        // Hard-coded scanners, reference src/setup.cu:300-372
        const float pi = kPiTrue;
        switch (scanner) {
            case 0:
                cg.r_f = 570.0f; cg.src_to_det = 1040.0f;
                cg.anode_angle = 7.0f * pi / 180.0f;
                cg.fan_angle_increment = 1.4083f / cg.src_to_det;
                cg.theta_cone = 2.0f * std::atan(7.5f * 1.3684f / cg.src_to_det);
                cg.central_channel = 335.25f; cg.n_proj_turn = 1160; cg.n_channels = 672;
                break;
            case 1:
                cg.r_f = 595.0f; cg.src_to_det = 1085.6f;
                cg.anode_angle = 7.0f * pi / 180.0f;
                cg.fan_angle_increment = 0.067864f * pi / 180.0f;
                cg.theta_cone = 2.0f * std::atan(7.5f * 1.2f / cg.r_f);
                cg.central_channel = 366.25f; cg.n_proj_turn = 1152; cg.n_channels = 736;
                break;
            case 2:
                cg.r_f = 570.0f; cg.src_to_det = 1040.0f;
                cg.anode_angle = 7.0f * pi / 180.0f;
                cg.fan_angle_increment = 0.07758621f * pi / 180.0f;
                cg.theta_cone = 2.0f * std::atan(7.5f * 1.2f / cg.r_f);
                cg.central_channel = 334.25f; cg.n_proj_turn = 1160; cg.n_channels = 672;
                break;
        }
    }
    // Derived quantities, reference src/setup.cu:291-297 (identical for every case)
    const float pi = kPiTrue;
    cg.n_proj_ffs = static_cast<size_t>(cg.n_proj_turn * std::pow(2, rp.phi_ffs) * std::pow(2, rp.z_ffs));
    cg.n_channels_oversampled = 2 * cg.n_channels;
    cg.n_rows = static_cast<unsigned int>(rp.n_rows);
    cg.n_rows_raw = static_cast<unsigned int>(rp.n_rows / std::pow(2, rp.z_ffs));
    cg.z_rot = rp.pitch_value;
    {
        float num = cg.fan_angle_increment * static_cast<float>(cg.n_channels) / 2;
        float den = 2.0f * pi / static_cast<float>(cg.n_proj_turn);
        cg.add_projections = static_cast<size_t>(num / den + 10);
    }
    cg.add_projections_ffs = static_cast<size_t>(cg.add_projections * std::pow(2, rp.z_ffs) * std::pow(2, rp.phi_ffs));
    cg.acq_fov = rp.acq_fov;
    if (rp.phi_ffs == 1) cg.central_channel = std::floor(cg.central_channel) + 0.375f;  // :377-380
    *out = cg;
    return CTBB_OK;
}

int ctbb_array_search(float key, const double *array, int numel_array, int search_type) {
    // reference src/setup.cu:806-825 (bounds are tested before the element)
    int idx = 0;
    switch (search_type) {
        case -1: while (idx < numel_array && key < array[idx]) idx++; break;
        case 0: while (idx < numel_array && key != array[idx]) idx++; break;
        case 1: while (idx < numel_array && key > array[idx]) idx++; break;
    }
    return idx;
}

}  // extern "C"

namespace ctbb {

// Tube angles and table positions for FileType 0 (binary): reference
// src/setup.cu:405-415 (float arithmetic kept).
void binary_angles_positions(const recon_params &rp, const ct_geom &cg, float *tube_angles,
                             double *table_positions) {
    for (int i = 0; i < rp.n_readings; i++) {
        tube_angles[i] = std::fmod(((360.0f / static_cast<float>(cg.n_proj_ffs)) * static_cast<float>(i) +
                                    rp.tube_start_angle),
                                   360.0f);
        float zr = cg.z_rot, nf = static_cast<float>(cg.n_proj_ffs);
        if (cg.table_direction == -1)
            table_positions[i] = (static_cast<float>(rp.n_readings) / nf) * zr - static_cast<float>(i) * zr / nf;
        else
            table_positions[i] = 0.0f + static_cast<float>(i) * zr / nf;
    }
}

struct PullRange {
    int start, end, n;
};

// Pull range around slice indices a,b: reference src/setup.cu:557-581 and
// 637-661 (identical arithmetic).
PullRange pull_range(const ct_geom &cg, int n_ffs, int idx_a, int idx_b) {
    int ppb = static_cast<int>(cg.n_proj_ffs / 2);
    int add_ffs = static_cast<int>(cg.add_projections_ffs);
    int s, e;
    if (idx_a > idx_b) {
        s = round_up_c(idx_b - ppb - add_ffs, n_ffs);
        e = round_up_c(idx_a + ppb + add_ffs, n_ffs);
    } else {
        s = round_up_c(idx_a - ppb - add_ffs, n_ffs);
        e = round_up_c(idx_b + ppb + add_ffs, n_ffs);
    }
    e += 256;
    int n = e - s;
    n = round_up_c(n, 128);
    e = s + n;
    return {s, e, n};
}

// reference src/setup.cu:385-603 without the raw-file reads (binary data) and
// optionally without the host allocations.
int plan_reconstruction(recon_metadata *mr, bool allocate) {
    ct_geom &cg = mr->cg;
    recon_params rp = mr->rp;
    if (rp.n_readings <= 1000) return CTBB_ERR_SHAPE;   // finalize reads table_positions[1000]
    if (rp.file_type != 0) return CTBB_ERR_UNSUPPORTED;

    std::free(mr->tube_angles);
    std::free(mr->table_positions);
    mr->tube_angles = static_cast<float *>(std::calloc(rp.n_readings, sizeof(float)));
    mr->table_positions = static_cast<double *>(std::calloc(rp.n_readings, sizeof(double)));
    if (!mr->tube_angles || !mr->table_positions) return CTBB_ERR_NOMEM;
    binary_angles_positions(rp, cg, mr->tube_angles, mr->table_positions);

    int n_ffs = static_cast<int>(std::pow(2, rp.z_ffs) * std::pow(2, rp.phi_ffs));
    int n_slices_block = kBlockSlices;
    int recon_direction = recon_direction_of(rp.start_pos, rp.end_pos);
    if (rp.n_slices != 0) rp.end_pos = rp.start_pos + (rp.n_slices - 1) * rp.slice_thickness;

    float recon_start_pos = rp.start_pos - recon_direction * rp.slice_thickness;
    float recon_end_pos = rp.end_pos + recon_direction * rp.slice_thickness;
    int n_slices_requested =
        static_cast<int>(std::floor(std::fabs(recon_end_pos - recon_start_pos) / rp.coll_slicewidth)) + 1;
    int n_slices_recon = (n_slices_requested - 1) + (n_slices_block - (n_slices_requested - 1) % n_slices_block);
    recon_end_pos = recon_start_pos + recon_direction * (n_slices_recon - 1) * rp.coll_slicewidth;
    int n_blocks = n_slices_recon / n_slices_block;

    const double *tp = mr->table_positions;
    int array_direction = direction_of(tp[0], tp[100]);
    int idx_slice_start = ctbb_array_search(recon_start_pos, tp, rp.n_readings, array_direction);
    int idx_slice_end = ctbb_array_search(recon_end_pos, tp, rp.n_readings, array_direction);

    mr->ri.data_begin_pos = static_cast<float>(tp[0]);
    mr->ri.data_end_pos = static_cast<float>(tp[rp.n_readings - 1]);
    float projection_padding =
        cg.z_rot * static_cast<float>(cg.n_proj_ffs / 2 + cg.add_projections_ffs + 256) / static_cast<float>(cg.n_proj_ffs);
    float allowed_begin = mr->ri.data_begin_pos + array_direction * projection_padding;
    float allowed_end = mr->ri.data_end_pos - array_direction * projection_padding;
    mr->ri.allowed_begin = allowed_begin;
    mr->ri.allowed_end = allowed_end;

    auto outside = [&](float p) {
        return (p > allowed_begin && p > allowed_end) || (p < allowed_begin && p < allowed_end);
    };
    if (outside(rp.start_pos) || outside(rp.end_pos)) return CTBB_ERR_RANGE;

    PullRange pr = pull_range(cg, n_ffs, idx_slice_start, idx_slice_end);

    cg.table_direction = array_direction;
    mr->rp.end_pos = rp.end_pos;
    mr->ri.n_ffs = n_ffs;
    mr->ri.n_slices_requested = n_slices_requested;
    mr->ri.n_slices_recon = n_slices_recon;
    mr->ri.n_slices_block = n_slices_block;
    mr->ri.n_blocks = n_blocks;
    mr->ri.idx_slice_start = idx_slice_start;
    mr->ri.idx_slice_end = idx_slice_end;
    mr->ri.recon_start_pos = recon_start_pos;
    mr->ri.recon_end_pos = recon_end_pos;
    mr->ri.idx_pull_start = static_cast<size_t>(pr.start);
    mr->ri.idx_pull_end = static_cast<size_t>(pr.end);
    mr->ri.n_proj_pull = static_cast<size_t>(pr.n);

    if (allocate) {
        std::free(mr->ctd.raw);
        std::free(mr->ctd.rebin);
        std::free(mr->ctd.image);
        size_t n_raw = cg.n_channels * cg.n_rows_raw * static_cast<size_t>(pr.n);
        size_t n_reb = cg.n_channels_oversampled * cg.n_rows *
                       (static_cast<size_t>(pr.n) - 2 * cg.add_projections_ffs) / n_ffs;
        size_t n_img = static_cast<size_t>(rp.nx) * rp.ny * n_slices_recon;
        mr->ctd.raw = static_cast<float *>(std::calloc(n_raw, sizeof(float)));
        mr->ctd.rebin = static_cast<float *>(std::calloc(n_reb, sizeof(float)));
        mr->ctd.image = static_cast<float *>(std::calloc(n_img, sizeof(float)));
        if (!mr->ctd.raw || !mr->ctd.rebin || !mr->ctd.image) return CTBB_ERR_NOMEM;
    }
    return CTBB_OK;
}

// reference src/setup.cu:605-679 for block `block_idx` (0-based).
void plan_block(recon_metadata *mr, int block_idx) {
    const recon_info &ri = mr->ri;
    const recon_params &rp = mr->rp;
    const ct_geom &cg = mr->cg;
    int n_ffs = static_cast<int>(std::pow(2, rp.z_ffs) * std::pow(2, rp.phi_ffs));
    int recon_direction = recon_direction_of(rp.start_pos, rp.end_pos);
    float block_slice_start = ri.recon_start_pos +
        recon_direction * block_idx * rp.coll_slicewidth * static_cast<float>(ri.n_slices_block);
    float block_slice_end = block_slice_start +
        static_cast<float>(recon_direction) * (static_cast<float>(ri.n_slices_block) - 1.0f) * rp.coll_slicewidth;
    const double *tp = mr->table_positions;
    int array_direction = direction_of(tp[0], tp[100]);
    int ia = ctbb_array_search(block_slice_start, tp, rp.n_readings, array_direction);
    int ib = ctbb_array_search(block_slice_end, tp, rp.n_readings, array_direction);
    PullRange pr = pull_range(cg, n_ffs, ia, ib);
    mr->ri.cb.block_slice_start = block_slice_start;
    mr->ri.cb.block_slice_end = block_slice_end;
    mr->ri.cb.idx_block_slice_start = ia;
    mr->ri.cb.idx_block_slice_end = ib;
    mr->ri.idx_pull_start = static_cast<size_t>(pr.start);
    mr->ri.idx_pull_end = static_cast<size_t>(pr.end);
    mr->ri.n_proj_pull = static_cast<size_t>(pr.n);
    mr->ri.cb.block_idx = block_idx + 1;
}

}  // namespace ctbb

extern "C" {

int ctbb_configure_reconstruction(struct recon_metadata *mr) {
    if (!mr) return CTBB_ERR_ARG;
    char full[4096 + 255 + 2] = {0};
    std::snprintf(full, sizeof(full), "%s/%s", mr->rp.raw_data_dir, mr->rp.raw_data_file);
    if (!file_exists(full)) return CTBB_ERR_IO;
    return ctbb::plan_reconstruction(mr, true);
}

int ctbb_plan_reconstruction(struct recon_metadata *mr) {
    if (!mr) return CTBB_ERR_ARG;
    return ctbb::plan_reconstruction(mr, false);
}

int ctbb_update_block_info(struct recon_metadata *mr) {
    if (!mr || !mr->table_positions) return CTBB_ERR_ARG;
    ctbb::plan_block(mr, mr->ri.cb.block_idx);
    const ct_geom &cg = mr->cg;
    std::free(mr->ctd.raw);
    std::free(mr->ctd.rebin);
    size_t n = mr->ri.n_proj_pull;
    size_t n_raw = cg.n_channels * cg.n_rows_raw * n;
    size_t n_reb = cg.n_channels_oversampled * cg.n_rows * (n - 2 * cg.add_projections_ffs) / mr->ri.n_ffs;
    mr->ctd.raw = static_cast<float *>(std::calloc(n_raw, sizeof(float)));
    mr->ctd.rebin = static_cast<float *>(std::calloc(n_reb, sizeof(float)));
    if (!mr->ctd.raw || !mr->ctd.rebin) return CTBB_ERR_NOMEM;
    return CTBB_OK;
}

int ctbb_plan_block(struct recon_metadata *mr, int block_idx) {
    if (!mr || !mr->table_positions || block_idx < 0) return CTBB_ERR_ARG;
    ctbb::plan_block(mr, block_idx);
    return CTBB_OK;
}

// FileType 0 reader: RawOffset header bytes, then float32 frames of
// n_channels*n_rows_raw samples, channel fastest.  Frames outside the file
// read as zeros (reference src/setup.cu:693-700; the reader itself lives in
// the un-vendored libctbb_read).
int ctbb_extract_projections(struct recon_metadata *mr) {
    if (!mr || !mr->ctd.raw) return CTBB_ERR_ARG;
    if (mr->rp.file_type != 0) return CTBB_ERR_UNSUPPORTED;
    char full[4096 + 255 + 2] = {0};
    std::snprintf(full, sizeof(full), "%s/%s", mr->rp.raw_data_dir, mr->rp.raw_data_file);
    FILE *f = std::fopen(full, "rb");
    if (!f) return CTBB_ERR_IO;
    const size_t frame = mr->cg.n_channels * mr->cg.n_rows_raw;
    long long start = static_cast<long long>(static_cast<int>(mr->ri.idx_pull_start));
    for (size_t i = 0; i < mr->ri.n_proj_pull; i++) {
        float *dst = mr->ctd.raw + i * frame;
        long long idx = start + static_cast<long long>(i);
        size_t got = 0;
        if (idx >= 0) {
            long long off = static_cast<long long>(mr->rp.raw_data_offset) +
                            idx * static_cast<long long>(frame * sizeof(float));
            if (fseeko(f, off, SEEK_SET) == 0) got = std::fread(dst, sizeof(float), frame, f);
        }
        if (got < frame) std::memset(dst + got, 0, (frame - got) * sizeof(float));
    }
    std::fclose(f);
    return CTBB_OK;
}

int ctbb_n_slices_final(const struct recon_metadata *mr) {
    if (!mr) return 0;
    // reference src/finalize_image_stack.cu:76
    return static_cast<int>(std::floor(std::fabs(mr->rp.end_pos - mr->rp.start_pos) / mr->rp.slice_thickness)) + 1;
}

int ctbb_finish_and_cleanup(struct recon_metadata *mr) {
    if (!mr) return CTBB_ERR_ARG;
    int status = CTBB_OK;
    if (mr->ctd.final_image_stack) {
        size_t n = static_cast<size_t>(mr->rp.nx) * mr->rp.ny * ctbb_n_slices_final(mr);
        char full[4096 + 255 + 2] = {0};
        std::snprintf(full, sizeof(full), "%s/%s", mr->rp.output_dir, mr->rp.output_file);
        FILE *f = std::fopen(full, "w");
        if (!f) status = CTBB_ERR_IO;
        else {
            if (std::fwrite(mr->ctd.final_image_stack, sizeof(float), n, f) != n) status = CTBB_ERR_IO;
            std::fclose(f);
        }
    }
    std::free(mr->ctd.rebin); mr->ctd.rebin = nullptr;
    std::free(mr->ctd.image); mr->ctd.image = nullptr;
    std::free(mr->ctd.raw); mr->ctd.raw = nullptr;
    std::free(mr->ctd.final_image_stack); mr->ctd.final_image_stack = nullptr;
    std::free(mr->tube_angles); mr->tube_angles = nullptr;
    std::free(mr->table_positions); mr->table_positions = nullptr;
    return status;
}

int ctbb_block_sizes(const struct recon_metadata *mr, ctbb_sizes *o) {
    if (!mr || !o) return CTBB_ERR_ARG;
    const ct_geom &cg = mr->cg;
    const recon_info &ri = mr->ri;
    if (ri.n_ffs == 0) return CTBB_ERR_ARG;
    std::memset(o, 0, sizeof(*o));
    o->raw_pull = ri.n_proj_pull * cg.n_rows_raw * cg.n_channels;
    size_t n_views = ri.n_proj_pull / ri.n_ffs;
    o->n_proj_out = n_views > 2 * cg.add_projections ? n_views - 2 * cg.add_projections : 0;
    o->rebin = o->n_proj_out * cg.n_rows * cg.n_channels_oversampled;
    o->image_block = static_cast<size_t>(ri.n_slices_block) * mr->rp.nx * mr->rp.ny;
    o->image_stack = ri.n_slices_recon * mr->rp.nx * mr->rp.ny;
    o->n_slices_final = ctbb_n_slices_final(mr);
    o->final_stack = static_cast<size_t>(o->n_slices_final) * mr->rp.nx * mr->rp.ny;
    // reference src/backproject.cu:39
    o->n_half_turns = static_cast<int>((ri.n_proj_pull / ri.n_ffs - 2 * cg.add_projections) / (cg.n_proj_turn / 2));
    o->voxel_updates_block = static_cast<double>(mr->rp.nx) * mr->rp.ny * ri.n_slices_block *
                             static_cast<double>(cg.n_proj_turn / 2) * o->n_half_turns;
    return CTBB_OK;
}

}  // extern "C"
