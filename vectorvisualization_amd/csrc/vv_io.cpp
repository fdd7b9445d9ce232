// vv_io.cpp — the reference's on-disk formats: DAT/RAW volumes (single and
// time-dependent), the raw noise file and its gradient cache, and 8-bit PNG
// for transfer functions and LIC filter kernels.  Host code; the loaders
// hand the decoded data to the same upload paths as the in-memory API.
//
//   DAT/RAW        DatFile::parseDatFile / readRawData   VV/reader.cpp:81-315
//   vector volume  VectorDataSet::loadData               VV/dataset.cpp:97-176
//   scalar volume  VolumeDataSet::loadData               VV/dataset.cpp:836-929
//   noise file     NoiseDataSet::loadRawData             VV/dataset.cpp:1347-1389
//   gradient cache loadGradients / saveGradients         VV/gradient.cpp:89-187
//   PNG            pngRead / pngWrite (libpng)           VV/imageUtils.cpp:19-136
//   filter PNG     LICFilter::loadData                   VV/dataset.cpp:1420-1467
//   TF PNGs        TransferEdit::setTFFileNames/saveTF/loadTF  VV/transferEdit.cpp:113-334
//
// PNG is decoded and encoded here over zlib (inflate/deflate, crc32); only
// what the reference's pngRead accepts is read: 8 bits per channel, gray,
// gray+alpha, RGB or RGBA, not interlaced.
#include "vv_ctx.hpp"

#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cstdint>
#include <cstdio>
#include <memory>

using namespace vv_host;

namespace {

struct File {
    FILE* f = nullptr;
    File(const char* path, const char* mode) : f(std::fopen(path, mode)) {}
    ~File()
    {
        if (f) std::fclose(f);
    }
    explicit operator bool() const { return f != nullptr; }
};

bool read_all(const char* path, std::vector<uint8_t>& out)
{
    File fp(path, "rb");
    if (!fp) return false;
    if (std::fseek(fp.f, 0, SEEK_END) != 0) return false;
    const long n = std::ftell(fp.f);
    if (n < 0 || std::fseek(fp.f, 0, SEEK_SET) != 0) return false;
    out.resize((size_t)n);
    return n == 0 || std::fread(out.data(), 1, (size_t)n, fp.f) == (size_t)n;
}

int raw_type_size(int t)
{
    switch (t) {
    case VV_RAW_UCHAR: return 1;
    case VV_RAW_USHORT: return 2;
    case VV_RAW_FLOAT: return 4;
    default: return 0;
    }
}

// the raw file name of a timestep: the ObjectFileName is a printf pattern
bool raw_name(const char* pattern, int t, char* out, size_t cap)
{
    const int n = std::snprintf(out, cap, pattern, t);
    return n >= 0 && (size_t)n < cap;
}

bool file_exists(const char* path)
{
    File fp(path, "rb");
    return (bool)fp;
}

// ObjectFileName may only hold one %d-style conversion (the timestep): any
// other conversion would read arguments that are not there.
bool pattern_ok(const char* p)
{
    int conv = 0;
    for (const char* c = p; *c; ++c) {
        if (*c != '%') continue;
        ++c;
        if (*c == '%') continue;
        while (*c && std::strchr("0123456789-+ #", *c)) ++c;
        if (*c != 'd' && *c != 'i') return false;
        ++conv;
    }
    return conv <= 1;
}

// ---- PNG --------------------------------------------------------------------
const uint8_t kPngSig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};

uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

void put_be32(std::vector<uint8_t>& v, uint32_t x)
{
    v.push_back((uint8_t)(x >> 24));
    v.push_back((uint8_t)(x >> 16));
    v.push_back((uint8_t)(x >> 8));
    v.push_back((uint8_t)x);
}

int png_channels(int color_type)
{
    switch (color_type) {
    case 0: return 1;  // gray
    case 4: return 2;  // gray + alpha
    case 2: return 3;  // RGB
    case 6: return 4;  // RGBA
    default: return 0; // palette and unknown: not read (as pngRead)
    }
}

uint8_t paeth(int a, int b, int c)
{
    const int p = a + b - c;
    const int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return (uint8_t)a;
    if (pb <= pc) return (uint8_t)b;
    return (uint8_t)c;
}

struct Png {
    int w = 0, h = 0, ch = 0;
    std::vector<uint8_t> px;  // rows top to bottom, as stored in the file
};

// header only (decode = false) or the whole image
vv_status png_decode(const char* path, Png& img, bool decode)
{
    std::vector<uint8_t> buf;
    if (!read_all(path, buf)) return fail(nullptr, VV_EIO, "cannot read PNG file \"%s\"", path);
    if (buf.size() < 8 || std::memcmp(buf.data(), kPngSig, 8) != 0)
        return fail(nullptr, VV_EIO, "\"%s\" is not a PNG file", path);
    size_t pos = 8;
    bool have_hdr = false, have_end = false;
    std::vector<uint8_t> idat;
    while (pos + 12 <= buf.size()) {
        const uint32_t len = be32(&buf[pos]);
        if (len > buf.size() - pos - 12) return fail(nullptr, VV_EIO, "PNG \"%s\": truncated chunk", path);
        const uint8_t* type = &buf[pos + 4];
        const uint8_t* data = &buf[pos + 8];
        const uint32_t crc = be32(&buf[pos + 8 + len]);
        char tname[5];
        for (int k = 0; k < 4; ++k) tname[k] = std::isalpha(type[k]) ? (char)type[k] : '?';
        tname[4] = 0;
        if ((uint32_t)crc32(crc32(0L, Z_NULL, 0), type, 4 + len) != crc)
            return fail(nullptr, VV_EIO, "PNG \"%s\": CRC mismatch in chunk %s", path, tname);
        if (!std::memcmp(type, "IHDR", 4)) {
            if (len != 13) return fail(nullptr, VV_EIO, "PNG \"%s\": bad IHDR", path);
            const uint32_t w = be32(data), h = be32(data + 4);
            const int depth = data[8], color = data[9], comp = data[10], filt = data[11], inter = data[12];
            if (depth != 8)
                return fail(nullptr, VV_EUNSUPPORTED, "PNG \"%s\": %d bits per channel (only 8 are read)", path,
                            depth);
            if (!png_channels(color))
                return fail(nullptr, VV_EUNSUPPORTED,
                            "PNG \"%s\": color type %d (only gray, gray+alpha, RGB, RGBA)", path, color);
            if (comp != 0 || filt != 0) return fail(nullptr, VV_EIO, "PNG \"%s\": unknown method", path);
            if (inter != 0) return fail(nullptr, VV_EUNSUPPORTED, "PNG \"%s\": interlaced", path);
            if (w == 0 || h == 0 || w > (1u << 16) || h > (1u << 16) ||
                (uint64_t)w * h * png_channels(color) > (1ull << 28))
                return fail(nullptr, VV_EUNSUPPORTED, "PNG \"%s\": size %ux%u (at most 2^28 bytes of pixels)",
                            path, w, h);
            img.w = (int)w;
            img.h = (int)h;
            img.ch = png_channels(color);
            have_hdr = true;
            if (!decode) return VV_OK;
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), data, data + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            have_end = true;
            break;
        } else if (!(type[0] & 0x20)) {
            return fail(nullptr, VV_EUNSUPPORTED, "PNG \"%s\": critical chunk %s", path, tname);
        }
        pos += 12 + len;
    }
    if (!have_hdr || !have_end) return fail(nullptr, VV_EIO, "PNG \"%s\": missing IHDR or IEND", path);
    const size_t stride = (size_t)img.w * img.ch;
    std::vector<uint8_t> raw((stride + 1) * img.h);
    z_stream zs;
    std::memset(&zs, 0, sizeof(zs));
    if (inflateInit(&zs) != Z_OK) return fail(nullptr, VV_ENOMEM, "zlib inflateInit failed");
    zs.next_in = idat.data();
    zs.avail_in = (uInt)idat.size();
    zs.next_out = raw.data();
    zs.avail_out = (uInt)raw.size();
    const int zr = inflate(&zs, Z_FINISH);
    const size_t got = raw.size() - zs.avail_out;
    inflateEnd(&zs);
    if (zr != Z_STREAM_END || got != raw.size())
        return fail(nullptr, VV_EIO, "PNG \"%s\": image data does not inflate to %dx%dx%d", path, img.w, img.h,
                    img.ch);
    img.px.resize(stride * img.h);
    const int bpp = img.ch;  // bytes per pixel at 8 bits
    for (int y = 0; y < img.h; ++y) {
        const uint8_t f = raw[y * (stride + 1)];
        const uint8_t* src = &raw[y * (stride + 1) + 1];
        uint8_t* dst = &img.px[y * stride];
        const uint8_t* up = y ? &img.px[(y - 1) * stride] : nullptr;
        for (size_t i = 0; i < stride; ++i) {
            const int a = i >= (size_t)bpp ? dst[i - bpp] : 0;
            const int b = up ? up[i] : 0;
            const int c = (up && i >= (size_t)bpp) ? up[i - bpp] : 0;
            int v;
            switch (f) {
            case 0: v = src[i]; break;
            case 1: v = src[i] + a; break;
            case 2: v = src[i] + b; break;
            case 3: v = src[i] + ((a + b) >> 1); break;
            case 4: v = src[i] + paeth(a, b, c); break;
            default: return fail(nullptr, VV_EIO, "PNG \"%s\": row filter %d", path, f);
            }
            dst[i] = (uint8_t)v;
        }
    }
    return VV_OK;
}

void put_chunk(std::vector<uint8_t>& out, const char* type, const uint8_t* data, size_t len)
{
    put_be32(out, (uint32_t)len);
    const size_t at = out.size();
    out.insert(out.end(), type, type + 4);
    if (len) out.insert(out.end(), data, data + len);
    put_be32(out, (uint32_t)crc32(crc32(0L, Z_NULL, 0), &out[at], (uInt)(4 + len)));
}

vv_status png_encode(const char* path, const uint8_t* px, int w, int h, int ch, bool flip)
{
    static const uint8_t color_of[5] = {0, 0, 4, 2, 6};
    const size_t stride = (size_t)w * ch;
    std::vector<uint8_t> raw((stride + 1) * h);
    for (int y = 0; y < h; ++y) {
        const int sy = flip ? h - 1 - y : y;
        raw[y * (stride + 1)] = 0;  // filter: none
        std::memcpy(&raw[y * (stride + 1) + 1], px + (size_t)sy * stride, stride);
    }
    uLongf zlen = compressBound((uLong)raw.size());
    std::vector<uint8_t> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) != Z_OK)
        return fail(nullptr, VV_ENOMEM, "zlib compress failed");
    std::vector<uint8_t> out(kPngSig, kPngSig + 8);
    uint8_t hdr[13];
    const uint32_t W = (uint32_t)w, H = (uint32_t)h;
    for (int i = 0; i < 4; ++i) {
        hdr[i] = (uint8_t)(W >> (24 - 8 * i));
        hdr[4 + i] = (uint8_t)(H >> (24 - 8 * i));
    }
    hdr[8] = 8;
    hdr[9] = color_of[ch];
    hdr[10] = hdr[11] = hdr[12] = 0;
    put_chunk(out, "IHDR", hdr, 13);
    put_chunk(out, "IDAT", z.data(), zlen);
    put_chunk(out, "IEND", nullptr, 0);
    File fp(path, "wb");
    if (!fp || std::fwrite(out.data(), 1, out.size(), fp.f) != out.size())
        return fail(nullptr, VV_EIO, "cannot write PNG file \"%s\"", path);
    return VV_OK;
}

// TransferEdit::setTFFileNames: "<base>_rgba<ext>" and "<base>_alpha<ext>",
// or "<name>_rgba.png" / "<name>_alpha.png" when there is no extension
void tf_names(const char* name, std::string& rgba, std::string& alpha)
{
    const char* dot = std::strrchr(name, '.');
    if (!dot) {
        rgba = std::string(name) + "_rgba.png";
        alpha = std::string(name) + "_alpha.png";
    } else {
        const std::string stem(name, (size_t)(dot - name));
        rgba = stem + "_rgba" + dot;
        alpha = stem + "_alpha" + dot;
    }
}

}  // namespace

extern "C" {

vv_status vv_dat_parse(const char* dat_path, vv_dat_info* info)
{
    return guarded(nullptr, [&]() -> vv_status {
        if (!dat_path || !info) return fail(nullptr, VV_EINVAL, "NULL argument");
        std::memset(info, 0, sizeof(*info));
        info->spacing[0] = info->spacing[1] = info->spacing[2] = 1.0f;
        info->components = 1;
        File fp(dat_path, "rb");
        if (!fp) return fail(nullptr, VV_EIO, "cannot open DAT file \"%s\"", dat_path);
        char line[1024];
        bool have_raw = false, have_res = false;
        while (std::fgets(line, sizeof(line), fp.f)) {
            if (line[0] == '#') continue;
            const char* colon = std::strchr(line, ':');
            // keys are matched anywhere in the line, in this order (reader.cpp:110-210)
            if (std::strstr(line, "ObjectFileName")) {
                char name[1024];
                if (!colon || std::sscanf(colon + 1, "%1023s", name) != 1)
                    return fail(nullptr, VV_EIO, "DAT \"%s\": bad ObjectFileName line", dat_path);
                std::snprintf(info->raw_pattern, sizeof(info->raw_pattern), "%s", name);
                have_raw = true;
            } else if (std::strstr(line, "Resolution")) {
                if (!colon || std::sscanf(colon + 1, "%d %d %d", &info->dims[0], &info->dims[1], &info->dims[2]) != 3)
                    return fail(nullptr, VV_EIO, "DAT \"%s\": bad Resolution line", dat_path);
                have_res = true;
            } else if (std::strstr(line, "SliceThickness")) {
                if (!colon ||
                    std::sscanf(colon + 1, "%f %f %f", &info->spacing[0], &info->spacing[1], &info->spacing[2]) != 3)
                    return fail(nullptr, VV_EIO, "DAT \"%s\": bad SliceThickness line", dat_path);
            } else if (std::strstr(line, "Format")) {
                const char* cp;
                if ((cp = std::strstr(line, "UCHAR")))
                    info->raw_type = VV_RAW_UCHAR;
                else if ((cp = std::strstr(line, "USHORT")))
                    info->raw_type = VV_RAW_USHORT;
                else if ((cp = std::strstr(line, "FLOAT")))
                    info->raw_type = VV_RAW_FLOAT;
                else
                    return fail(nullptr, VV_EUNSUPPORTED, "DAT \"%s\": Format other than UCHAR, USHORT, FLOAT",
                                dat_path);
                // components: the first integer after the type name (parseDataDim), else 1
                while (*cp && !std::isdigit((unsigned char)*cp)) ++cp;
                int dim = 1;
                if (*cp && std::sscanf(cp, "%d", &dim) != 1) dim = 1;
                info->components = dim;
            } else if (std::strstr(line, "TimeDependent")) {
                if (!colon || std::sscanf(colon + 1, "%d %d", &info->t_begin, &info->t_end) != 2)
                    return fail(nullptr, VV_EIO, "DAT \"%s\": bad TimeDependent line", dat_path);
                if (info->t_begin < 0 || info->t_end < info->t_begin)
                    return fail(nullptr, VV_EINVAL, "DAT \"%s\": illegal time steps %d-%d", dat_path, info->t_begin,
                                info->t_end);
            }
            // other lines are skipped
        }
        if (!have_raw || !have_res || info->raw_type == VV_RAW_NONE)
            return fail(nullptr, VV_EIO, "DAT \"%s\": ObjectFileName, Resolution and Format are required", dat_path);
        if (info->components < 1 || info->components > 4)
            return fail(nullptr, VV_EUNSUPPORTED, "DAT \"%s\": %d components per voxel", dat_path, info->components);
        for (int a = 0; a < 3; ++a)
            if (info->dims[a] < 1 || info->dims[a] > 4096 || !(info->spacing[a] > 0.0f))
                return fail(nullptr, VV_EINVAL, "DAT \"%s\": bad resolution or slice thickness", dat_path);
        if (!pattern_ok(info->raw_pattern))
            return fail(nullptr, VV_EINVAL, "DAT \"%s\": ObjectFileName may hold one %%d (the time step) only",
                        dat_path);
        // the RAW file as given, else next to the DAT file (reader.cpp:216-240)
        char name[1200];
        if (!raw_name(info->raw_pattern, info->t_begin, name, sizeof(name)) || !file_exists(name)) {
            const char* slash = std::strrchr(dat_path, '/');
            const char* bslash = std::strrchr(dat_path, '\\');
            const char* sep = std::max(slash, bslash);
            if (!sep) return fail(nullptr, VV_EIO, "DAT \"%s\": RAW file \"%s\" not found", dat_path, name);
            std::string dir(dat_path, (size_t)(sep - dat_path + 1));
            const std::string joined = dir + info->raw_pattern;
            if (joined.size() >= sizeof(info->raw_pattern))
                return fail(nullptr, VV_EINVAL, "DAT \"%s\": RAW path too long", dat_path);
            std::snprintf(info->raw_pattern, sizeof(info->raw_pattern), "%s", joined.c_str());
            if (!raw_name(info->raw_pattern, info->t_begin, name, sizeof(name)) || !file_exists(name))
                return fail(nullptr, VV_EIO, "DAT \"%s\": cannot open RAW file \"%s\"", dat_path, name);
        }
        for (int t = info->t_begin + 1; t <= info->t_end; ++t)
            if (!raw_name(info->raw_pattern, t, name, sizeof(name)) || !file_exists(name))
                return fail(nullptr, VV_EIO, "DAT \"%s\": cannot open RAW file \"%s\" of time step %d", dat_path, name,
                            t);
        return VV_OK;
    });
}

vv_status vv_dat_read(const vv_dat_info* info, int timestep, void* out, size_t out_bytes)
{
    return guarded(nullptr, [&]() -> vv_status {
        if (!info || !out) return fail(nullptr, VV_EINVAL, "NULL argument");
        if (timestep < info->t_begin || timestep > info->t_end)
            return fail(nullptr, VV_EINVAL, "time step %d outside %d..%d", timestep, info->t_begin, info->t_end);
        const size_t need = (size_t)raw_type_size(info->raw_type) * info->components * info->dims[0] * info->dims[1] *
                            info->dims[2];
        if (!need || out_bytes < need)
            return fail(nullptr, VV_EINVAL, "output holds %zu bytes, %zu needed", out_bytes, need);
        char name[1200];
        if (!raw_name(info->raw_pattern, timestep, name, sizeof(name)))
            return fail(nullptr, VV_EINVAL, "RAW file name too long");
        File fp(name, "rb");
        if (!fp) return fail(nullptr, VV_EIO, "cannot open RAW file \"%s\"", name);
        if (std::fread(out, 1, need, fp.f) != need)
            return fail(nullptr, VV_EIO, "RAW file \"%s\" holds fewer than %zu bytes", name, need);
        return VV_OK;
    });
}

vv_status vv_dat_read_field(const vv_dat_info* info, int timestep, float* out3)
{
    return guarded(nullptr, [&]() -> vv_status {
        if (!info || !out3) return fail(nullptr, VV_EINVAL, "NULL argument");
        if (info->components != 3)
            return fail(nullptr, VV_EINVAL, "not a vector data set (%d components)", info->components);
        if (info->raw_type != VV_RAW_UCHAR && info->raw_type != VV_RAW_FLOAT)
            return fail(nullptr, VV_EUNSUPPORTED, "vector data must be UCHAR or FLOAT (VV/dataset.cpp:131-137)");
        const size_t n = (size_t)3 * info->dims[0] * info->dims[1] * info->dims[2];
        if (info->raw_type == VV_RAW_FLOAT) return vv_dat_read(info, timestep, out3, n * sizeof(float));
        std::vector<uint8_t> u(n);
        vv_status s = vv_dat_read(info, timestep, u.data(), n);
        if (s) return s;
        for (size_t i = 0; i < n; ++i) out3[i] = (float)((int)u[i] - 128);  // UCHAR components are offset by 128
        return VV_OK;
    });
}

vv_status vv_load_field_dat(vv_ctx* ctx, const char* dat_path, int timestep, int next_timestep, float lerp_t,
                            int storage)
{
    return guarded(ctx, [&]() -> vv_status {
        if (!ctx) return VV_EINVAL;
        vv_dat_info info;
        vv_status s = vv_dat_parse(dat_path, &info);
        if (s) return fail(ctx, s, "%s", g_create_error.c_str());
        const size_t n = (size_t)3 * info.dims[0] * info.dims[1] * info.dims[2];
        std::vector<float> v0(n), v1;
        s = vv_dat_read_field(&info, timestep, v0.data());
        if (!s && next_timestep != timestep && lerp_t != 0.0f) {
            v1.resize(n);
            s = vv_dat_read_field(&info, next_timestep, v1.data());
        }
        if (s) return fail(ctx, s, "%s", g_create_error.c_str());
        return vv_upload_field(ctx, v0.data(), v1.empty() ? nullptr : v1.data(), v1.empty() ? 0.0f : lerp_t, info.dims,
                               info.spacing, storage, 0);
    });
}

vv_status vv_load_scalar_dat(vv_ctx* ctx, const char* dat_path)
{
    return guarded(ctx, [&]() -> vv_status {
        if (!ctx) return VV_EINVAL;
        vv_dat_info info;
        vv_status s = vv_dat_parse(dat_path, &info);
        if (s) return fail(ctx, s, "%s", g_create_error.c_str());
        if (info.components != 1) return fail(ctx, VV_EINVAL, "\"%s\" is not a scalar data set", dat_path);
        const size_t n = (size_t)info.dims[0] * info.dims[1] * info.dims[2];
        std::vector<uint8_t> u(n);
        if (info.raw_type == VV_RAW_UCHAR) {
            s = vv_dat_read(&info, info.t_begin, u.data(), n);
        } else if (info.raw_type == VV_RAW_FLOAT) {
            // a float volume in an 8-bit luminance texture: GL's UNORM8 conversion
            std::vector<float> f(n);
            s = vv_dat_read(&info, info.t_begin, f.data(), n * sizeof(float));
            for (size_t i = 0; !s && i < n; ++i) {
                const float c = std::isnan(f[i]) ? 0.0f : std::min(std::max(f[i], 0.0f), 1.0f);
                u[i] = (uint8_t)std::lrint(c * 255.0f);
            }
        } else {
            return fail(ctx, VV_EUNSUPPORTED, "scalar data must be UCHAR or FLOAT (VV/dataset.cpp:866-872)");
        }
        if (s) return fail(ctx, s, "%s", g_create_error.c_str());
        return vv_upload_scalar(ctx, u.data(), info.dims, 0);
    });
}

vv_status vv_noise_file_read(const char* path, int dims[3], uint8_t* out, size_t out_bytes)
{
    return guarded(nullptr, [&]() -> vv_status {
        if (!path || !dims) return fail(nullptr, VV_EINVAL, "NULL argument");
        File fp(path, "rb");
        if (!fp) return fail(nullptr, VV_EIO, "cannot open noise file \"%s\"", path);
        int32_t hdr[3];
        if (std::fread(hdr, 4, 3, fp.f) != 3) return fail(nullptr, VV_EIO, "noise file \"%s\": no header", path);
        for (int a = 0; a < 3; ++a) {
            if (hdr[a] < 2 || hdr[a] > 4096) return fail(nullptr, VV_EIO, "noise file \"%s\": bad size", path);
            dims[a] = hdr[a];
        }
        if (!out) return VV_OK;
        const size_t n = (size_t)dims[0] * dims[1] * dims[2];
        if (out_bytes < n) return fail(nullptr, VV_EINVAL, "output holds %zu bytes, %zu needed", out_bytes, n);
        if (std::fread(out, 1, n, fp.f) != n)
            return fail(nullptr, VV_EIO, "noise file \"%s\" holds fewer than %zu voxels", path, n);
        return VV_OK;
    });
}

vv_status vv_noise_file_write(const char* path, const uint8_t* noise, const int dims[3])
{
    return guarded(nullptr, [&]() -> vv_status {
        if (!path || !noise || !dims) return fail(nullptr, VV_EINVAL, "NULL argument");
        if (!dims_ok(dims)) return fail(nullptr, VV_EINVAL, "noise dims %s", kDimsRule);
        File fp(path, "wb");
        const int32_t hdr[3] = {dims[0], dims[1], dims[2]};
        const size_t n = (size_t)dims[0] * dims[1] * dims[2];
        if (!fp || std::fwrite(hdr, 4, 3, fp.f) != 3 || std::fwrite(noise, 1, n, fp.f) != n)
            return fail(nullptr, VV_EIO, "cannot write noise file \"%s\"", path);
        return VV_OK;
    });
}

vv_status vv_load_noise_file(vv_ctx* ctx, const char* path, int with_gradients, int use_gradient_cache)
{
    return guarded(ctx, [&]() -> vv_status {
        if (!ctx) return VV_EINVAL;
        int dims[3];
        vv_status s = vv_noise_file_read(path, dims, nullptr, 0);
        if (s) return fail(ctx, s, "%s", g_create_error.c_str());
        const size_t n = (size_t)dims[0] * dims[1] * dims[2];
        std::vector<uint8_t> noise(n);
        s = vv_noise_file_read(path, dims, noise.data(), n);
        if (s) return fail(ctx, s, "%s", g_create_error.c_str());
        if (!with_gradients) return vv_upload_noise(ctx, noise.data(), dims, 0, nullptr, 0);
        // gradients: "<noise file>.grd" holds 3 bytes per voxel (createGradientsFileName)
        const std::string grd = std::string(path) + ".grd";
        std::vector<uint8_t> g;
        if (use_gradient_cache && read_all(grd.c_str(), g) && g.size() == 3 * n)
            return vv_upload_noise(ctx, noise.data(), dims, 1, g.data(), 0);
        s = vv_upload_noise(ctx, noise.data(), dims, 1, nullptr, 0);
        if (s || !use_gradient_cache) return s;
        std::vector<uint8_t> rgba(4 * n);
        s = vv_download_noise(ctx, rgba.data(), 0);
        if (s) return s;
        g.resize(3 * n);
        for (size_t i = 0; i < n; ++i)
            for (int c = 0; c < 3; ++c) g[3 * i + c] = rgba[4 * i + c];
        File fp(grd.c_str(), "wb");
        if (!fp || std::fwrite(g.data(), 1, g.size(), fp.f) != g.size())
            return fail(ctx, VV_EIO, "cannot write gradient cache \"%s\"", grd.c_str());
        return VV_OK;
    });
}

vv_status vv_load_noise_or_generate(vv_ctx* ctx, const char* path, int with_gradients, int use_gradient_cache,
                                    uint64_t seed, int* generated)
{
    return guarded(ctx, [&]() -> vv_status {
        if (!ctx) return VV_EINVAL;
        if (generated) *generated = 0;
        int dims[3];
        const vv_status r = path ? vv_noise_file_read(path, dims, nullptr, 0) : VV_EIO;
        if (r == VV_OK) {
            // the size check reads the whole file; a short file falls back too
            const size_t n = (size_t)dims[0] * dims[1] * dims[2];
            std::vector<uint8_t> probe(n);
            if (vv_noise_file_read(path, dims, probe.data(), n) == VV_OK)
                return vv_load_noise_file(ctx, path, with_gradients, use_gradient_cache);
        }
        // "NoiseData:  Using random noise 256x256x256": 255 where
        // floor(0.6 u + 0.5) = 1, i.e. u >= 5/6
        const int d256[3] = {256, 256, 256};
        const vv_status s = vv_generate_noise(ctx, seed, VV_NOISE_BINARY, 1.0 / 6.0, d256, with_gradients);
        if (s == VV_OK && generated) *generated = 1;
        return s;
    });
}

vv_status vv_png_read(const char* path, uint8_t* out, size_t out_bytes, int* width, int* height, int* channels)
{
    return guarded(nullptr, [&]() -> vv_status {
        if (!path) return fail(nullptr, VV_EINVAL, "NULL path");
        Png img;
        vv_status s = png_decode(path, img, out != nullptr);
        if (s) return s;
        if (width) *width = img.w;
        if (height) *height = img.h;
        if (channels) *channels = img.ch;
        if (!out) return VV_OK;
        if (out_bytes < img.px.size())
            return fail(nullptr, VV_EINVAL, "output holds %zu bytes, %zu needed", out_bytes, img.px.size());
        std::memcpy(out, img.px.data(), img.px.size());
        return VV_OK;
    });
}

vv_status vv_png_write(const char* path, const uint8_t* pixels, int width, int height, int channels,
                       int flip_rows)
{
    return guarded(nullptr, [&]() -> vv_status {
        if (!path || !pixels) return fail(nullptr, VV_EINVAL, "NULL argument");
        if (width < 1 || height < 1 || width > (1 << 16) || height > (1 << 16) || channels < 1 || channels > 4)
            return fail(nullptr, VV_EINVAL, "bad PNG size or channel count");
        return png_encode(path, pixels, width, height, channels, flip_rows != 0);
    });
}

vv_status vv_image_write_png(const char* path, const float* rgba, int width, int height)
{
    return guarded(nullptr, [&]() -> vv_status {
        if (!path || !rgba) return fail(nullptr, VV_EINVAL, "NULL argument");
        if (width < 1 || height < 1 || width > (1 << 16) || height > (1 << 16))
            return fail(nullptr, VV_EINVAL, "bad image size");
        const size_t n = (size_t)width * height * 4;
        std::vector<uint8_t> px(n);
        for (size_t i = 0; i < n; ++i) {
            const float c = std::isnan(rgba[i]) ? 0.0f : std::min(std::max(rgba[i], 0.0f), 1.0f);
            px[i] = (uint8_t)std::lrint(c * 255.0f);  // UNORM8, as a glReadPixels of the frame
        }
        // images are stored bottom row first; PNG rows run top to bottom
        return png_encode(path, px.data(), width, height, 4, true);
    });
}

vv_status vv_load_filter_png(vv_ctx* ctx, const char* path)
{
    return guarded(ctx, [&]() -> vv_status {
        if (!ctx) return VV_EINVAL;
        Png img;
        vv_status s = png_decode(path, img, true);
        if (s) return fail(ctx, s, "%s", g_create_error.c_str());
        // first row, first channel, centred in the next power of two (VV/dataset.cpp:1445-1457)
        int w = 1;
        while (w < img.w) w <<= 1;
        if (w > 4096) return fail(ctx, VV_EUNSUPPORTED, "filter \"%s\" wider than 4096", path);
        std::vector<uint8_t> lut((size_t)w, 0);
        const int shift = (w - img.w) / 2;
        for (int i = 0; i < img.w; ++i) lut[i + shift] = img.px[(size_t)i * img.ch];
        return vv_set_filter(ctx, lut.data(), w, 0.0f);
    });
}

vv_status vv_tf_load_png(const char* name, uint8_t tf5[1280])
{
    return guarded(nullptr, [&]() -> vv_status {
        if (!name || !tf5) return fail(nullptr, VV_EINVAL, "NULL argument");
        std::string rgba_name, alpha_name;
        tf_names(name, rgba_name, alpha_name);
        // loadTF runs loadRGBATF() || loadAlphaOpacTF() (VV/transferEdit.cpp:238):
        // the alpha/opacity file is read only when the RGBA file cannot be.
        Png img;
        if (png_decode(rgba_name.c_str(), img, true) == VV_OK && (size_t)img.w * img.h >= 256) {
            for (int i = 0; i < 256; ++i) {
                const uint8_t* p = &img.px[(size_t)i * img.ch];
                if (img.ch < 3) {
                    tf5[5 * i] = tf5[5 * i + 1] = tf5[5 * i + 2] = p[0];
                    if (img.ch == 2) tf5[5 * i + 3] = p[1];
                } else {
                    tf5[5 * i] = p[0];
                    tf5[5 * i + 1] = p[1];
                    tf5[5 * i + 2] = p[2];
                    if (img.ch == 4) tf5[5 * i + 3] = p[3];
                }
            }
            return VV_OK;
        }
        Png al;
        vv_status s = png_decode(alpha_name.c_str(), al, true);
        if (s)
            return fail(nullptr, s, "no transfer function in \"%s\" or \"%s\"", rgba_name.c_str(),
                        alpha_name.c_str());
        if ((size_t)al.w * al.h < 256)
            return fail(nullptr, VV_EIO, "\"%s\" holds fewer than 256 entries", alpha_name.c_str());
        for (int i = 0; i < 256; ++i) {
            const uint8_t* p = &al.px[(size_t)i * al.ch];
            tf5[5 * i + 3] = p[0];
            if (al.ch == 2) tf5[5 * i + 4] = p[1];
        }
        return VV_OK;
    });
}

vv_status vv_tf_save_png(const char* name, const uint8_t tf5[1280])
{
    return guarded(nullptr, [&]() -> vv_status {
        if (!name || !tf5) return fail(nullptr, VV_EINVAL, "NULL argument");
        std::string rgba_name, alpha_name;
        tf_names(name, rgba_name, alpha_name);
        // 10 identical rows of 256 entries (VV/transferEdit.cpp:176-214)
        std::vector<uint8_t> px(4 * 256 * 10);
        for (int j = 0; j < 10; ++j)
            for (int i = 0; i < 256; ++i)
                for (int c = 0; c < 4; ++c) px[4 * (j * 256 + i) + c] = tf5[5 * i + c];
        vv_status s = png_encode(rgba_name.c_str(), px.data(), 256, 10, 4, false);
        if (s) return s;
        for (int j = 0; j < 10; ++j)
            for (int i = 0; i < 256; ++i) {
                px[2 * (j * 256 + i)] = tf5[5 * i + 3];
                px[2 * (j * 256 + i) + 1] = tf5[5 * i + 4];
            }
        return png_encode(alpha_name.c_str(), px.data(), 256, 10, 2, false);
    });
}

}  // extern "C"
