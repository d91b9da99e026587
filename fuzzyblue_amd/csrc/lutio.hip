// lutio.hip — LUT files: fb_atmosphere_save / fb_atmosphere_load.
//
// The reference's dump example reads the three LUTs back into linear
// buffers (examples/dump.rs:110-126, 175-193: x fastest, then y, then the
// 3D texture's depth) and writes each one as an OpenEXR image
// (examples/dump.rs:331-367): the image is width x height, a 2D LUT has
// channels R, G, B, A, and each depth slice z of the scattering LUT is a
// layer of channels "{z}.R", "{z}.G", "{z}.B", "{z}.A"; transmittance and
// irradiance are 32-bit float, scattering half.
//
// Here the same images are written as single-part scanline OpenEXR 2.0
// files with no compression (the layout is the published OpenEXR file
// format: magic and version, a header of named attributes, a table of
// scanline offsets, then per scanline its y, its byte count and, per
// channel in the sorted channel list, that channel's row of samples).  Two
// extra header attributes make a set loadable: "fuzzyblueLut" (string: which
// LUT) and "fuzzyblueParameters" (the fb_parameters struct's bytes, host
// byte order).  fb_atmosphere_load reads such files back into a finished
// atmosphere; it accepts exactly the subset the writer produces.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "runtime.hpp"

using namespace fbrt;

namespace {

constexpr int32_t kExrMagic = 20000630;
constexpr int32_t kExrVersion = 2;  // single-part scanline, no long names
constexpr int32_t kPixelHalf = 1, kPixelFloat = 2;
const char* const kLutNames[3] = {"transmittance", "irradiance", "scattering"};

uint16_t f32_to_f16(float f) {
    _Float16 h = static_cast<_Float16>(f);  // round to nearest even
    uint16_t u;
    memcpy(&u, &h, 2);
    return u;
}
float f16_to_f32(uint16_t u) {
    _Float16 h;
    memcpy(&h, &u, 2);
    return static_cast<float>(h);
}

// A channel of the image: its name and the (layer, component) it holds.
struct Channel {
    std::string name;
    uint32_t layer, comp;  // comp 0..3 = R, G, B, A
};

// examples/dump.rs:358-367: "R".."A" for one layer, "{layer}.R".. otherwise;
// OpenEXR lists channels sorted by name (byte order).
std::vector<Channel> channels_of(uint32_t depth) {
    static const char rgba[4] = {'R', 'G', 'B', 'A'};
    std::vector<Channel> ch;
    for (uint32_t z = 0; z < depth; ++z)
        for (uint32_t c = 0; c < 4; ++c)
            ch.push_back(Channel{depth == 1 ? std::string(1, rgba[c])
                                            : std::to_string(z) + "." + rgba[c],
                                 z, c});
    std::sort(ch.begin(), ch.end(),
              [](const Channel& a, const Channel& b) { return a.name < b.name; });
    return ch;
}

struct Out {
    std::vector<uint8_t> b;
    template <typename T>
    void put(const T& v) {
        const auto* p = reinterpret_cast<const uint8_t*>(&v);
        b.insert(b.end(), p, p + sizeof(T));
    }
    void str(const std::string& s) {
        b.insert(b.end(), s.begin(), s.end());
        b.push_back(0);
    }
    void attr(const char* name, const char* type, const void* v, size_t n) {
        str(name);
        str(type);
        put(int32_t(n));
        const auto* p = static_cast<const uint8_t*>(v);
        b.insert(b.end(), p, p + n);
    }
};

struct In {
    const uint8_t* p;
    size_t n, pos = 0;
    bool ok = true;
    template <typename T>
    T get() {
        T v{};
        if (pos + sizeof(T) > n) {
            ok = false;
            return v;
        }
        memcpy(&v, p + pos, sizeof(T));
        pos += sizeof(T);
        return v;
    }
    std::string str() {
        std::string s;
        while (pos < n && p[pos]) s.push_back(char(p[pos++]));
        if (pos >= n) ok = false;
        ++pos;
        return s;
    }
};

bool read_file(const char* path, std::vector<uint8_t>& out) {
    FILE* f = fopen(path, "rb");
    if (!f) return false;
    uint8_t buf[1 << 16];
    size_t k;
    while ((k = fread(buf, 1, sizeof(buf), f)) > 0) out.insert(out.end(), buf, buf + k);
    const bool err = ferror(f) != 0;
    fclose(f);
    return !err;
}

// A parsed image: header fields this module needs and the samples as float
// RGBA texels in the linear layout.
struct Image {
    std::string lut;
    fb_parameters params{};
    bool have_params = false;
    int32_t pixel_type = 0;
    uint32_t w = 0, h = 0, depth = 0;
    std::vector<float> rgba;      // w*h*depth*4 (FLOAT channels)
    std::vector<uint16_t> rgba16;  // w*h*depth*4 (HALF channels)
};

fb_status parse_exr(const char* path, Image& img) {
    std::vector<uint8_t> file;
    if (!read_file(path, file)) {
        set_error("cannot read %s", path);
        return FB_EINVAL;
    }
    In in{file.data(), file.size()};
    if (in.get<int32_t>() != kExrMagic || in.get<int32_t>() != kExrVersion) {
        set_error("%s is not a single-part scanline OpenEXR 2 file", path);
        return FB_EINVAL;
    }
    std::vector<std::pair<std::string, int32_t>> chans;
    int32_t box[4] = {0, 0, -1, -1};
    bool have_box = false;
    while (in.ok) {
        const std::string name = in.str();
        if (name.empty()) break;
        const std::string type = in.str();
        const int32_t size = in.get<int32_t>();
        if (!in.ok || size < 0 || in.pos + size_t(size) > in.n) {
            set_error("%s: truncated header", path);
            return FB_EINVAL;
        }
        const size_t end = in.pos + size_t(size);
        if (name == "channels" && type == "chlist") {
            while (in.pos < end) {
                const std::string cn = in.str();
                if (cn.empty()) break;
                const int32_t pt = in.get<int32_t>();
                (void)in.get<uint32_t>();  // pLinear + reserved
                const int32_t xs = in.get<int32_t>(), ys = in.get<int32_t>();
                if (xs != 1 || ys != 1) {
                    set_error("%s: subsampled channel %s", path, cn.c_str());
                    return FB_EINVAL;
                }
                chans.emplace_back(cn, pt);
            }
        } else if (name == "compression" && type == "compression") {
            if (size != 1 || file[in.pos] != 0) {
                set_error("%s: compressed OpenEXR is not supported (written uncompressed)", path);
                return FB_EINVAL;
            }
        } else if (name == "lineOrder" && type == "lineOrder") {
            if (size != 1 || file[in.pos] != 0) {
                set_error("%s: line order must be increasing y", path);
                return FB_EINVAL;
            }
        } else if (name == "dataWindow" && type == "box2i" && size == 16) {
            memcpy(box, file.data() + in.pos, 16);
            have_box = true;
        } else if (name == "fuzzyblueLut" && type == "string") {
            img.lut.assign(reinterpret_cast<const char*>(file.data() + in.pos), size_t(size));
        } else if (name == "fuzzyblueParameters" && size == int32_t(sizeof(fb_parameters))) {
            memcpy(&img.params, file.data() + in.pos, sizeof(fb_parameters));
            img.have_params = true;
        }
        in.pos = end;
    }
    if (!in.ok || !have_box || chans.empty() || box[0] != 0 || box[1] != 0 || box[2] < 0 ||
        box[3] < 0) {
        set_error("%s: missing or unsupported channels / dataWindow", path);
        return FB_EINVAL;
    }
    img.w = uint32_t(box[2]) + 1;
    img.h = uint32_t(box[3]) + 1;
    if (chans.size() % 4 != 0) {
        set_error("%s: %zu channels is not RGBA layers", path, chans.size());
        return FB_EINVAL;
    }
    img.depth = uint32_t(chans.size() / 4);
    const std::vector<Channel> want = channels_of(img.depth);
    img.pixel_type = chans[0].second;
    for (size_t i = 0; i < chans.size(); ++i)
        if (chans[i].first != want[i].name || chans[i].second != img.pixel_type ||
            (img.pixel_type != kPixelHalf && img.pixel_type != kPixelFloat)) {
            set_error("%s: channel %zu is %s (type %d); expected %s", path, i,
                      chans[i].first.c_str(), chans[i].second, want[i].name.c_str());
            return FB_EINVAL;
        }
    const size_t bps = img.pixel_type == kPixelHalf ? 2 : 4;
    const size_t texels = size_t(img.w) * img.h * img.depth;
    const size_t line = want.size() * img.w * bps;
    in.pos += size_t(img.h) * 8;  // scanline offset table (lines are in order)
    if (img.pixel_type == kPixelHalf)
        img.rgba16.assign(texels * 4, 0);
    else
        img.rgba.assign(texels * 4, 0.0f);
    for (uint32_t y = 0; y < img.h; ++y) {
        const int32_t yy = in.get<int32_t>(), bytes = in.get<int32_t>();
        if (!in.ok || yy != int32_t(y) || size_t(bytes) != line || in.pos + line > in.n) {
            set_error("%s: bad scanline %u", path, y);
            return FB_EINVAL;
        }
        for (const Channel& c : want) {
            const size_t base = (size_t(c.layer) * img.h + y) * img.w;
            for (uint32_t x = 0; x < img.w; ++x) {
                const size_t dst = (base + x) * 4 + c.comp;
                if (img.pixel_type == kPixelHalf)
                    memcpy(&img.rgba16[dst], file.data() + in.pos, 2);
                else
                    memcpy(&img.rgba[dst], file.data() + in.pos, 4);
                in.pos += bps;
            }
        }
    }
    return FB_OK;
}

}  // namespace

namespace fbrt {
// runtime.hip: a finished atmosphere's T, E, S (and events) without scratch.
fb_status alloc_atmosphere(fb_builder* b, const fb_parameters& p, fb_atmosphere** out);
}  // namespace fbrt

extern "C" {

fb_status fb_atmosphere_save(const fb_atmosphere* a, fb_lut which, uint32_t format,
                             const char* path) {
    if (!a || !path) {
        set_error("NULL argument");
        return FB_EINVAL;
    }
    if (which != FB_LUT_TRANSMITTANCE && which != FB_LUT_IRRADIANCE &&
        which != FB_LUT_SCATTERING) {
        set_error("LUT %d is not an atmosphere output", int(which));
        return FB_EINVAL;
    }
    if (format != FB_FORMAT_F32 && format != FB_FORMAT_F16) {
        set_error("unknown format %u", format);
        return FB_EINVAL;
    }
    // The reference's files: float transmittance and irradiance, half
    // scattering (examples/dump.rs:110-126); FB_FORMAT_F16 applies to the
    // scattering LUT only.
    const bool half = which == FB_LUT_SCATTERING && format == FB_FORMAT_F16;
    fb_lut_view v;
    fb_status st = fb_atmosphere_lut(a, which, &v);
    if (st != FB_OK) return st;
    const size_t texels = size_t(v.width) * v.height * v.depth;
    std::vector<float> f32;
    std::vector<uint16_t> f16;
    if (half) {
        f16.resize(texels * 4);
        st = fb_atmosphere_download(a, which, FB_FORMAT_F16, f16.data(), f16.size() * 2);
    } else {
        f32.resize(texels * 4);
        st = fb_atmosphere_download(a, which, FB_FORMAT_F32, f32.data(), f32.size() * 4);
    }
    if (st != FB_OK) return st;

    const std::vector<Channel> ch = channels_of(v.depth);
    Out o;
    o.put(kExrMagic);
    o.put(kExrVersion);
    {
        Out cl;
        for (const Channel& c : ch) {
            cl.str(c.name);
            cl.put(int32_t(half ? kPixelHalf : kPixelFloat));
            cl.put(uint32_t(0));  // pLinear 0, reserved
            cl.put(int32_t(1));
            cl.put(int32_t(1));
        }
        cl.b.push_back(0);
        o.attr("channels", "chlist", cl.b.data(), cl.b.size());
    }
    const uint8_t none = 0;
    o.attr("compression", "compression", &none, 1);
    const int32_t box[4] = {0, 0, int32_t(v.width) - 1, int32_t(v.height) - 1};
    o.attr("dataWindow", "box2i", box, 16);
    o.attr("displayWindow", "box2i", box, 16);
    o.attr("lineOrder", "lineOrder", &none, 1);
    const float one = 1.0f, center[2] = {0.0f, 0.0f};
    o.attr("pixelAspectRatio", "float", &one, 4);
    o.attr("screenWindowCenter", "v2f", center, 8);
    o.attr("screenWindowWidth", "float", &one, 4);
    const char* lut = kLutNames[int(which)];
    o.attr("fuzzyblueLut", "string", lut, strlen(lut));
    o.attr("fuzzyblueParameters", "fbParameters", &a->params, sizeof(fb_parameters));
    o.b.push_back(0);  // end of header
    const size_t bps = half ? 2 : 4;
    const size_t line = ch.size() * v.width * bps;
    const size_t table = o.b.size();
    for (uint32_t y = 0; y < v.height; ++y)
        o.put(uint64_t(table + size_t(v.height) * 8 + size_t(y) * (8 + line)));
    for (uint32_t y = 0; y < v.height; ++y) {
        o.put(int32_t(y));
        o.put(int32_t(line));
        for (const Channel& c : ch) {
            const size_t base = (size_t(c.layer) * v.height + y) * v.width;
            for (uint32_t x = 0; x < v.width; ++x) {
                const size_t src = (base + x) * 4 + c.comp;
                if (half)
                    o.put(f16[src]);
                else
                    o.put(f32[src]);
            }
        }
    }
    FILE* f = fopen(path, "wb");
    if (!f) {
        set_error("cannot create %s", path);
        return FB_EINVAL;
    }
    const bool ok = fwrite(o.b.data(), 1, o.b.size(), f) == o.b.size();
    if (fclose(f) != 0 || !ok) {
        set_error("cannot write %s", path);
        return FB_EINVAL;
    }
    return FB_OK;
}

fb_status fb_atmosphere_load(fb_builder* b, const char* transmittance_path,
                             const char* irradiance_path, const char* scattering_path,
                             fb_atmosphere** out) {
    if (!b || !transmittance_path || !irradiance_path || !scattering_path || !out) {
        set_error("NULL argument");
        return FB_EINVAL;
    }
    *out = nullptr;
    const char* paths[3] = {transmittance_path, irradiance_path, scattering_path};
    Image img[3];
    for (int k = 0; k < 3; ++k) {
        fb_status st = parse_exr(paths[k], img[k]);
        if (st != FB_OK) return st;
        if (img[k].lut != kLutNames[k] || !img[k].have_params) {
            set_error("%s is not a fuzzyblue %s LUT file", paths[k], kLutNames[k]);
            return FB_EINVAL;
        }
        if (memcmp(&img[k].params, &img[0].params, sizeof(fb_parameters)) != 0) {
            set_error("%s was saved from another atmosphere than %s", paths[k], paths[0]);
            return FB_EINVAL;
        }
        if (k < 2 && img[k].pixel_type != kPixelFloat) {
            set_error("%s: %s must be 32-bit float", paths[k], kLutNames[k]);
            return FB_EINVAL;
        }
    }
    const fb_parameters& p = img[0].params;
    fb_status st = fb_parameters_validate(&p);
    if (st != FB_OK) return st;
    uint32_t te[2], ie[2], se[3];
    fb_parameters_extents(&p, te, ie, se);
    const uint32_t dims[3][3] = {{te[0], te[1], 1}, {ie[0], ie[1], 1}, {se[0], se[1], se[2]}};
    for (int k = 0; k < 3; ++k)
        if (img[k].w != dims[k][0] || img[k].h != dims[k][1] || img[k].depth != dims[k][2]) {
            set_error("%s: %ux%ux%u does not match the parameters' %ux%ux%u", paths[k],
                      img[k].w, img[k].h, img[k].depth, dims[k][0], dims[k][1], dims[k][2]);
            return FB_EINVAL;
        }
    fb_atmosphere* a = nullptr;
    st = alloc_atmosphere(b, p, &a);
    if (st != FB_OK) return st;
    auto fail = [&](hipError_t e, const char* what) {
        fb_atmosphere_destroy(a);
        return hip_fail(e, what);
    };
    hipError_t e = hipMemcpy(a->T.p, img[0].rgba.data(), a->T.bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(a->E.p, img[1].rgba.data(), a->E.bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(e, "LUT upload");
    // The scattering LUT in the set's storage format: binary16 widens
    // exactly, float32 rounds to nearest even.
    const bool stored_f16 = (p.flags & FB_FLAG_F16_STORAGE) != 0;
    const Image& s = img[2];
    if (stored_f16 && s.pixel_type == kPixelHalf) {
        e = hipMemcpy(a->S.p, s.rgba16.data(), a->S.bytes, hipMemcpyHostToDevice);
    } else if (stored_f16) {
        std::vector<uint16_t> h(s.rgba.size());
        for (size_t i = 0; i < h.size(); ++i) h[i] = f32_to_f16(s.rgba[i]);
        e = hipMemcpy(a->S.p, h.data(), a->S.bytes, hipMemcpyHostToDevice);
    } else if (s.pixel_type == kPixelHalf) {
        std::vector<float> f(s.rgba16.size());
        for (size_t i = 0; i < f.size(); ++i) f[i] = f16_to_f32(s.rgba16[i]);
        e = hipMemcpy(a->S.p, f.data(), a->S.bytes, hipMemcpyHostToDevice);
    } else {
        e = hipMemcpy(a->S.p, s.rgba.data(), a->S.bytes, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hipEventRecord(a->ready, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(a->ready);
    if (e != hipSuccess) return fail(e, "LUT upload");
    *out = a;
    return FB_OK;
}

}  // extern "C"
