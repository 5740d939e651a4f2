// Dataset readers for the depth input of the per-frame loop: the RAW (SLAMBench) format and the
// TUM RGB-D directory format, with ground-truth poses; host code.
//
// Reference: app/src/reader_base.cpp (nextDataImpl 536-572, readPose 408-477, mergeStatus 361-366),
// app/src/reader_raw.cpp (RAWReader 10-46, readResolution 74-87, nextDepth 91-122),
// app/src/reader_tum.cpp (image lists, ground truth, association, pose interpolation, TUMReader,
// nextPose, nextDepth), app/include/reader_tum.hpp:46-54 (constants).
//
// The reference decodes TUM's 16-bit PNG depth with OpenCV; here a small PNG decoder over zlib
// (non-interlaced greyscale, 8 or 16 bits, CRC-checked) does it.  Colour images are not read (the
// path is depth-only), so a TUM directory needs depth/, depth.txt and rgb.txt (for the
// association) but no rgb/ images.  Depth is returned either as stored (uint16 plus the
// conversion the reference applies, for the device-side conversion fused into
// se8_downsample_depth_u16) or converted on the host as the reference does.
#include "se8_internal.hpp"

#include <zlib.h>

#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace {

constexpr float kTumInverseScale = 1.0f / 5000.0f; // reader_tum.hpp:46
constexpr double kMaxMatchDist = 0.02;             // reader_tum.hpp:52
constexpr double kMaxInterpDist = 10.0 * kMaxMatchDist;

bool is_dir(const std::string& p)
{
    struct stat st;
    return stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

bool is_file(const std::string& p)
{
    struct stat st;
    return stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

// Split on spaces, dropping empty fields (str_utils::split_str with consecutive delimiters merged).
std::vector<std::string> split(const std::string& line)
{
    std::vector<std::string> out;
    std::string cur;
    for (char c : line) {
        if (c == ' ' || c == '\r') {
            if (!cur.empty())
                out.push_back(cur);
            cur.clear();
        }
        else
            cur.push_back(c);
    }
    if (!cur.empty())
        out.push_back(cur);
    return out;
}

bool parse_double(const std::string& s, double& v, bool allow_negative = true)
{
    if (s.empty())
        return false;
    char* end = nullptr;
    v = std::strtod(s.c_str(), &end);
    return end == s.c_str() + s.size() && (allow_negative || v >= 0.0);
}

// ---------------------------------------------------------------------------------------------
// PNG (ISO/IEC 15948): signature, chunks with CRC-32, zlib stream of filtered scanlines.
// ---------------------------------------------------------------------------------------------
uint32_t be32(const unsigned char* p)
{
    return ((uint32_t) p[0] << 24) | ((uint32_t) p[1] << 16) | ((uint32_t) p[2] << 8) | (uint32_t) p[3];
}

int paeth(int a, int b, int c)
{
    const int p = a + b - c;
    const int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc)
        return a;
    return pb <= pc ? b : c;
}

bool read_png_gray(const std::string& path, int& w, int& h, std::vector<uint16_t>& out, std::string& err)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    std::vector<unsigned char> buf((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    static const unsigned char sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    if (buf.size() < 8 || std::memcmp(buf.data(), sig, 8) != 0) {
        err = path + " is not a PNG file";
        return false;
    }
    int depth = 0, ctype = -1;
    std::vector<unsigned char> idat;
    size_t i = 8;
    bool have_ihdr = false, have_iend = false;
    while (i + 12 <= buf.size()) {
        const uint32_t len = be32(&buf[i]);
        if (i + 12 + (size_t) len > buf.size())
            break;
        const unsigned char* type = &buf[i + 4];
        const unsigned char* data = &buf[i + 8];
        const uint32_t crc = be32(&buf[i + 8 + len]);
        if ((uint32_t) crc32(crc32(0L, Z_NULL, 0), type, len + 4) != crc) {
            err = path + ": chunk CRC mismatch";
            return false;
        }
        if (std::memcmp(type, "IHDR", 4) == 0 && len == 13) {
            w = (int) be32(data);
            h = (int) be32(data + 4);
            depth = data[8];
            ctype = data[9];
            if (data[10] != 0 || data[11] != 0 || data[12] != 0) {
                err = path + ": interlaced or non-standard PNG";
                return false;
            }
            have_ihdr = true;
        }
        else if (std::memcmp(type, "IDAT", 4) == 0)
            idat.insert(idat.end(), data, data + len);
        else if (std::memcmp(type, "IEND", 4) == 0) {
            have_iend = true;
            break;
        }
        i += 12 + (size_t) len;
    }
    if (!have_ihdr || !have_iend || ctype != 0 || (depth != 8 && depth != 16) || w <= 0 || h <= 0) {
        err = path + ": only complete greyscale PNGs of 8 or 16 bits are read";
        return false;
    }
    const size_t bpp = (size_t) depth / 8;
    const size_t stride = (size_t) w * bpp;
    std::vector<unsigned char> raw((stride + 1) * (size_t) h);
    uLongf raw_len = (uLongf) raw.size();
    if (uncompress(raw.data(), &raw_len, idat.data(), (uLong) idat.size()) != Z_OK || raw_len != raw.size()) {
        err = path + ": corrupt image data";
        return false;
    }
    std::vector<unsigned char> prev(stride, 0), cur(stride);
    out.resize((size_t) w * h);
    for (int y = 0; y < h; y++) {
        const unsigned char* line = &raw[(size_t) y * (stride + 1)];
        const int filter = line[0];
        for (size_t x = 0; x < stride; x++) {
            const int a = x >= bpp ? cur[x - bpp] : 0, b = prev[x], c = x >= bpp ? prev[x - bpp] : 0;
            int v = line[1 + x];
            switch (filter) {
            case 0: break;
            case 1: v += a; break;
            case 2: v += b; break;
            case 3: v += (a + b) / 2; break;
            case 4: v += paeth(a, b, c); break;
            default:
                err = path + ": bad scanline filter";
                return false;
            }
            cur[x] = (unsigned char) v;
        }
        for (int x = 0; x < w; x++)
            out[(size_t) y * w + x] = depth == 16 ? (uint16_t) ((cur[2 * x] << 8) | cur[2 * x + 1]) : cur[x];
        std::swap(prev, cur);
    }
    return true;
}

// ---------------------------------------------------------------------------------------------
// poses
// ---------------------------------------------------------------------------------------------
struct Quat {
    float w, x, y, z;
};

// Eigen::Quaternion::toRotationMatrix, row-major into T (translation column left alone).
void quat_to_matrix(const Quat& q, float* T)
{
    const float tx = 2.f * q.x, ty = 2.f * q.y, tz = 2.f * q.z;
    const float twx = tx * q.w, twy = ty * q.w, twz = tz * q.w;
    const float txx = tx * q.x, txy = ty * q.x, txz = tz * q.x;
    const float tyy = ty * q.y, tyz = tz * q.y, tzz = tz * q.z;
    T[0] = 1.f - (tyy + tzz), T[1] = txy - twz, T[2] = txz + twy;
    T[4] = txy + twz, T[5] = 1.f - (txx + tzz), T[6] = tyz - twx;
    T[8] = txz - twy, T[9] = tyz + twx, T[10] = 1.f - (txx + tyy);
}

float quat_norm(const Quat& q)
{
    return std::sqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
}

// Spherical linear interpolation along the shorter arc; linear weights when the two orientations
// are within float epsilon of parallel (Eigen::QuaternionBase::slerp).
Quat slerp(const Quat& a, const Quat& b, float t)
{
    const float one = 1.f - 1.1920929e-07f;
    const float d = a.w * b.w + a.x * b.x + a.y * b.y + a.z * b.z;
    const float ad = std::fabs(d);
    float s0, s1;
    if (ad >= one) {
        s0 = 1.f - t;
        s1 = t;
    }
    else {
        const float theta = std::acos(ad);
        const float st = std::sin(theta);
        s0 = std::sin((1.f - t) * theta) / st;
        s1 = std::sin(t * theta) / st;
    }
    if (d < 0.f)
        s1 = -s1;
    return Quat{s0 * a.w + s1 * b.w, s0 * a.x + s1 * b.x, s0 * a.y + s1 * b.y, s0 * a.z + s1 * b.z};
}

void make_pose(const float p[3], const Quat& q, float* T)
{
    for (int k = 0; k < 16; k++)
        T[k] = (k % 5 == 0) ? 1.f : 0.f;
    quat_to_matrix(q, T);
    T[3] = p[0], T[7] = p[1], T[11] = p[2];
}

struct TumImage {
    double t;
    std::string file;
};

struct TumPose {
    double t = NAN;
    float p[3] = {0, 0, 0};
    Quat q{1, 0, 0, 0};
};

bool read_image_list(const std::string& path, std::vector<TumImage>& out, std::string& err)
{
    std::ifstream f(path);
    if (!f) {
        err = "cannot read " + path;
        return false;
    }
    for (std::string line; std::getline(f, line);) {
        if (!line.empty() && line[0] == '#')
            continue;
        const std::vector<std::string> c = split(line);
        double t;
        if (c.size() != 2 || !parse_double(c[0], t, false)) {
            err = path + ": each data line must be 'timestamp filename' with a non-negative timestamp (line " +
                  std::to_string(out.size() + 1) + ")";
            return false;
        }
        out.push_back({t, c[1]});
    }
    if (out.empty()) {
        err = "empty file " + path;
        return false;
    }
    return true;
}

bool read_tum_ground_truth(const std::string& path, std::vector<TumPose>& out, std::string& err)
{
    std::ifstream f(path);
    if (!f) {
        err = "cannot read ground truth file " + path;
        return false;
    }
    for (std::string line; std::getline(f, line);) {
        if (!line.empty() && line[0] == '#')
            continue;
        const std::vector<std::string> c = split(line);
        double v[8];
        bool ok = c.size() == 8 && parse_double(c[0], v[0], false);
        for (size_t k = 1; ok && k < 8; k++)
            ok = parse_double(c[k], v[k]);
        if (!ok) {
            err = path + ": each data line must be 'timestamp tx ty tz qx qy qz qw' (line " +
                  std::to_string(out.size() + 1) + ")";
            return false;
        }
        TumPose p;
        p.t = v[0];
        for (int k = 0; k < 3; k++)
            p.p[k] = std::stof(c[1 + k]);
        p.q = Quat{std::stof(c[7]), std::stof(c[4]), std::stof(c[5]), std::stof(c[6])};
        out.push_back(p);
    }
    if (out.empty()) {
        err = "empty ground truth file " + path;
        return false;
    }
    return true;
}

} // namespace

struct se8_reader {
    se8_reader_config cfg{};
    std::string seq, gt_path;
    int width = 0, height = 0;
    int64_t num_frames = 0;
    size_t frame = SIZE_MAX;
    int status = SE8_READER_OK;
    float scale = 1000.f; // RAW: millimetres, divided
    int divide = 1;
    // RAW
    std::ifstream raw;
    size_t depth_total = 0, colour_total = 0;
    std::ifstream gt;
    size_t gt_frame = SIZE_MAX;
    // TUM
    std::vector<std::string> depth_files;
    std::vector<TumPose> poses; // interpolated at the depth timestamps
};

namespace {

int merge_status(int a, int b)
{
    return a > b ? a : b;
}

bool good(int s)
{
    return s == SE8_READER_OK || s == SE8_READER_SKIP;
}

// Reader::readPose (reader_base.cpp:408-477): the frame-th data line's last 7 columns.
int read_pose_line(se8_reader* r, float* T)
{
    std::string line;
    while (true) {
        std::getline(r->gt, line);
        if (!r->gt.good())
            return SE8_READER_EOF;
        if (!line.empty() && line[0] == '#')
            continue;
        r->gt_frame++;
        if (r->gt_frame < r->frame)
            continue;
        const std::vector<std::string> c = split(line);
        if (c.size() < 7)
            return SE8_READER_ERROR;
        float v[7];
        bool finite = true;
        for (int k = 0; k < 7; k++) {
            double d;
            if (!parse_double(c[c.size() - 7 + k], d))
                return SE8_READER_ERROR;
            v[k] = std::stof(c[c.size() - 7 + k]);
            finite = finite && std::isfinite(v[k]);
        }
        if (!finite)
            return SE8_READER_SKIP;
        const Quat q{v[6], v[3], v[4], v[5]};
        if (std::fabs(quat_norm(q) - 1.f) > 1e-3f)
            return SE8_READER_SKIP;
        make_pose(v, q, T);
        return SE8_READER_OK;
    }
}

int next_pose(se8_reader* r, float* T)
{
    if (r->cfg.type == SE8_READER_TUM) {
        if (r->poses.empty() || r->frame >= r->poses.size())
            return SE8_READER_ERROR;
        const TumPose& p = r->poses[r->frame];
        if (std::fabs(quat_norm(p.q) - 1.f) > 1e-3f)
            return SE8_READER_SKIP;
        make_pose(p.p, p.q, T);
        return SE8_READER_OK;
    }
    if (!r->gt.is_open())
        return SE8_READER_ERROR;
    return read_pose_line(r, T);
}

bool read_resolution(std::ifstream& f, int& w, int& h)
{
    uint32_t v[2];
    if (!f.read(reinterpret_cast<char*>(v), sizeof(v)))
        return false;
    w = (int) v[0];
    h = (int) v[1];
    return true;
}

int next_depth_raw(se8_reader* r, uint16_t* depth)
{
    if (r->cfg.type == SE8_READER_RAW) {
        r->raw.clear();
        r->raw.seekg((std::streamoff) (r->frame * (r->depth_total + r->colour_total)));
        int w, h;
        if (!read_resolution(r->raw, w, h) || w != r->width || h != r->height)
            return SE8_READER_ERROR;
        if (!r->raw.read(reinterpret_cast<char*>(depth), sizeof(uint16_t) * (size_t) w * h))
            return SE8_READER_ERROR;
        return SE8_READER_OK;
    }
    if (r->frame >= r->depth_files.size())
        return SE8_READER_ERROR;
    int w, h;
    std::vector<uint16_t> img;
    std::string err;
    if (!read_png_gray(r->seq + "/" + r->depth_files[r->frame], w, h, img, err) || w != r->width ||
        h != r->height)
        return SE8_READER_ERROR;
    std::memcpy(depth, img.data(), sizeof(uint16_t) * img.size());
    return SE8_READER_OK;
}

int open_raw(se8_reader* r, std::string& err)
{
    r->raw.open(r->seq, std::ios::in | std::ios::binary);
    if (!r->raw.good()) {
        err = "could not read raw file " + r->seq;
        return SE8_ERR_IO;
    }
    int cw, ch;
    if (!read_resolution(r->raw, r->width, r->height) || r->width <= 0 || r->height <= 0) {
        err = "could not read the depth image size of " + r->seq;
        return SE8_ERR_IO;
    }
    r->depth_total = 8 + sizeof(uint16_t) * (size_t) r->width * r->height;
    r->raw.seekg((std::streamoff) r->depth_total);
    if (!read_resolution(r->raw, cw, ch)) {
        err = "could not read the colour image size of " + r->seq;
        return SE8_ERR_IO;
    }
    r->colour_total = 8 + 3 * (size_t) cw * ch;
    r->raw.clear();
    r->raw.seekg(0, std::ios::end);
    const size_t size = (size_t) r->raw.tellg();
    r->raw.seekg(0);
    r->num_frames = (int64_t) (size / (r->depth_total + r->colour_total));
    r->scale = 1000.f;
    r->divide = 1;
    if (!r->gt_path.empty()) {
        r->gt.open(r->gt_path);
        if (!r->gt.good()) {
            err = "could not read ground truth file " + r->gt_path;
            return SE8_ERR_IO;
        }
    }
    return SE8_OK;
}

int open_tum(se8_reader* r, std::string& err)
{
    r->scale = r->cfg.inverse_scale != 0.f ? r->cfg.inverse_scale : kTumInverseScale;
    r->divide = 0;
    if (!is_dir(r->seq) || !is_dir(r->seq + "/depth") || !is_file(r->seq + "/depth.txt") ||
        !is_file(r->seq + "/rgb.txt")) {
        err = "a TUM sequence path must be a directory with a depth subdirectory and depth.txt and rgb.txt files";
        return SE8_ERR_IO;
    }
    std::vector<TumImage> depth, rgb;
    if (!read_image_list(r->seq + "/depth.txt", depth, err) || !read_image_list(r->seq + "/rgb.txt", rgb, err))
        return SE8_ERR_IO;
    // associate each depth image with the RGB image of the closest timestamp (within 0.02 s)
    std::vector<TumImage> kept;
    for (const TumImage& d : depth) {
        double best = INFINITY;
        for (const TumImage& c : rgb)
            best = std::fmin(best, std::fabs(d.t - c.t));
        if (best <= kMaxMatchDist)
            kept.push_back(d);
    }
    if (!r->gt_path.empty()) {
        std::vector<TumPose> gt;
        if (!read_tum_ground_truth(r->gt_path, gt, err))
            return SE8_ERR_IO;
        std::vector<TumImage> with_pose;
        for (const TumImage& d : kept) {
            // the surrounding ground-truth poses (reader_tum.cpp, surrounding_poses)
            TumPose a, b;
            for (size_t k = 0; k + 1 < gt.size(); k++) {
                const TumPose& p = gt[k];
                const TumPose& n = gt[k + 1];
                if (p.t == d.t) {
                    a = b = p;
                    break;
                }
                else if (n.t == d.t)
                    a = b = n;
                else if (p.t < d.t && d.t < n.t) {
                    a = p;
                    b = n;
                    break;
                }
            }
            const double diff = b.t - a.t;
            if (std::isnan(diff) || std::fabs(diff) > kMaxInterpDist)
                continue;
            const double t = b.t != a.t ? (d.t - a.t) / diff : 0.0;
            TumPose p;
            p.t = d.t;
            for (int k = 0; k < 3; k++)
                p.p[k] = (float) (1.0f - t) * a.p[k] + (float) t * b.p[k];
            p.q = slerp(a.q, b.q, (float) t);
            r->poses.push_back(p);
            with_pose.push_back(d);
        }
        if (with_pose.empty()) {
            err = "could not associate any ground truth poses to depth images";
            return SE8_ERR_IO;
        }
        kept = with_pose;
    }
    for (const TumImage& d : kept)
        r->depth_files.push_back(d.file);
    r->num_frames = (int64_t) r->depth_files.size();
    if (!r->depth_files.empty()) {
        std::vector<uint16_t> img;
        if (!read_png_gray(r->seq + "/" + r->depth_files[0], r->width, r->height, img, err))
            return SE8_ERR_IO;
    }
    return SE8_OK;
}

} // namespace

extern "C" int se8_reader_create(const se8_reader_config* cfg, se8_reader** out)
{
    SE8_REQUIRE(cfg && out && cfg->sequence_path, "se8_reader_create: null argument");
    SE8_REQUIRE(cfg->type == SE8_READER_RAW || cfg->type == SE8_READER_TUM, "se8_reader_create: unknown reader type");
    *out = nullptr;
    se8_reader* r = new se8_reader;
    r->cfg = *cfg;
    r->seq = cfg->sequence_path;
    r->gt_path = cfg->ground_truth_file ? cfg->ground_truth_file : "";
    std::string err;
    const int rc = cfg->type == SE8_READER_RAW ? open_raw(r, err) : open_tum(r, err);
    if (rc != SE8_OK) {
        se8::set_error("se8_reader_create: " + err);
        delete r;
        return rc;
    }
    *out = r;
    return SE8_OK;
}

extern "C" int se8_reader_destroy(se8_reader* r)
{
    delete r;
    return SE8_OK;
}

extern "C" int se8_reader_info(const se8_reader* r, int* width, int* height, int64_t* num_frames,
                               int* has_ground_truth)
{
    SE8_REQUIRE(r, "se8_reader_info: null reader");
    if (width)
        *width = r->width;
    if (height)
        *height = r->height;
    if (num_frames)
        *num_frames = r->num_frames;
    if (has_ground_truth)
        *has_ground_truth = !r->gt_path.empty();
    return SE8_OK;
}

extern "C" int se8_reader_next_raw(se8_reader* r, uint16_t* depth, float* scale, int* divide, float T_WB[16],
                                   int* status)
{
    SE8_REQUIRE(r && depth && status, "se8_reader_next_raw: null argument");
    if (scale)
        *scale = r->scale;
    if (divide)
        *divide = r->divide;
    if (!good(r->status)) {
        *status = r->status;
        return SE8_OK;
    }
    r->frame++;
    r->status = next_depth_raw(r, depth);
    if (good(r->status) && T_WB)
        r->status = merge_status(next_pose(r, T_WB), r->status);
    *status = r->status;
    return SE8_OK;
}

extern "C" int se8_reader_next(se8_reader* r, float* depth, float T_WB[16], int* status)
{
    SE8_REQUIRE(r && depth && status, "se8_reader_next: null argument");
    std::vector<uint16_t> raw((size_t) r->width * r->height);
    float scale;
    int divide;
    const int rc = se8_reader_next_raw(r, raw.data(), &scale, &divide, T_WB, status);
    if (rc != SE8_OK || !good(*status))
        return rc;
    for (size_t i = 0; i < raw.size(); i++)
        depth[i] = divide ? (float) raw[i] / scale : (float) raw[i] * scale;
    return SE8_OK;
}

extern "C" int se8_reader_restart(se8_reader* r)
{
    SE8_REQUIRE(r, "se8_reader_restart: null reader");
    r->frame = SIZE_MAX;
    r->gt_frame = SIZE_MAX;
    r->status = SE8_READER_OK;
    if (r->gt.is_open()) {
        r->gt.clear();
        r->gt.seekg(0);
    }
    if (r->raw.is_open()) {
        r->raw.clear();
        r->raw.seekg(0);
    }
    return SE8_OK;
}
