// nd_image.cpp - height-map images for the @image formula (formulas/heightmap_primitive.wgsl).
//
// The reference decodes an image file with stb_image and prepares it on the host
// (height_maps.zig: decode, quantize, slopeCode, dilate); its shader samples the result
// (image_info / image_texel / image_sample, template.wgsl:1110-1193). Here the host keeps
// decoding (the pixels stbi_load_16 or stbi_loadf return cross the C ABI) and this file does
// everything after the pixel read, in the same f32 arithmetic:
//   - RGBA16: straight alpha -> (r a, g a, b a, a) in [0, 1]; RGB32F (HDR): RGB normalised
//     by the finite minimum and maximum over every channel, alpha 1;
//   - box downsampling by the smallest integer factor that brings both edges to 2048 or
//     less, then quantization to 16 bits;
//   - per texel the largest change toward its right / lower neighbours (slope), its maximum
//     over the image, and that slope dilated (window maximum) over 8 + 2 and 64 + 2 texels;
//   - 4 words per texel: RG and BA as 16-bit pairs, the two dilated slopes as 8-bit codes.
// Bound images live in one device buffer per context: 16 slot headers of 8 words (texel
// offset, width, height, per-channel slope bound) followed by the texels, the layout the
// shader reads.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "nd_ctx.h"

namespace {

constexpr uint32_t kMaxEdge = ND_IMAGE_MAX_EDGE;
constexpr uint32_t kRadiusNear = 8, kRadiusFar = 64, kDilateMargin = 2;
constexpr float kSlopeStepsPerOctave = 16.f, kSlopeCodeMax = 255.f, kSlopeCodeMaxLog2 = 1.f, kSlopeRoundUp = 1.001f;

struct Px { float c[4]; };

inline float q16(float v) {
    // quantize (height_maps.zig): round(clamp(v, 0, 1) * 65535) / 65535, round half away from 0
    return std::round(std::min(std::max(v, 0.f), 1.f) * 65535.f) / 65535.f;
}

// Window maximum of radius r along one line (stride apart): out[j] = max over
// [j - r, j + r] clipped to the line, by the van Herk / Gil-Werman forward and backward runs
// over blocks of 2r + 1.
void dilate_line(Px* data, size_t start, size_t stride, size_t len, size_t r, std::vector<Px>& fwd,
                 std::vector<Px>& bwd, std::vector<Px>& out) {
    const size_t k = 2 * r + 1;
    for (size_t i = 0; i < len; ++i) {
        const Px v = data[start + i * stride];
        if (i % k == 0) fwd[i] = v;
        else for (int ch = 0; ch < 4; ++ch) fwd[i].c[ch] = std::max(fwd[i - 1].c[ch], v.c[ch]);
    }
    for (size_t i = len; i-- > 0;) {
        const Px v = data[start + i * stride];
        if (i == len - 1 || (i + 1) % k == 0) bwd[i] = v;
        else for (int ch = 0; ch < 4; ++ch) bwd[i].c[ch] = std::max(bwd[i + 1].c[ch], v.c[ch]);
    }
    for (size_t j = 0; j < len; ++j) {
        const size_t lo = j > r ? j - r : 0, hi = std::min(j + r, len - 1);
        for (int ch = 0; ch < 4; ++ch) out[j].c[ch] = std::max(bwd[lo].c[ch], fwd[hi].c[ch]);
    }
    for (size_t j = 0; j < len; ++j) data[start + j * stride] = out[j];
}

void dilate(std::vector<Px>& d, size_t w, size_t h, size_t r) {
    const size_t n = std::max(w, h);
    std::vector<Px> a(n), b(n), o(n);
    for (size_t y = 0; y < h; ++y) dilate_line(d.data(), y * w, 1, w, r, a, b, o);
    for (size_t x = 0; x < w; ++x) dilate_line(d.data(), x, w, h, r, a, b, o);
}

uint32_t slope_code(float s) {
    if (!(s > 0.f)) return 0;
    const float steps = std::ceil(kSlopeStepsPerOctave * (std::log2(s * kSlopeRoundUp) - kSlopeCodeMaxLog2) + kSlopeCodeMax);
    return (uint32_t)std::min(std::max(steps, 1.f), kSlopeCodeMax);
}
uint32_t pack_slopes(const Px& s) {
    uint32_t o = 0;
    for (int ch = 0; ch < 4; ++ch) o |= slope_code(s.c[ch]) << (8 * ch);
    return o;
}

// The decoded pixels as straight floats (decode's first half).
bool to_pixels(nd_image_format fmt, const void* data, uint32_t w, uint32_t h, std::vector<Px>& px) {
    const size_t n = (size_t)w * h;
    px.resize(n);
    if (fmt == ND_IMAGE_RGBA16) {
        const uint16_t* s = (const uint16_t*)data;
        const float inv = 1.f / 65535.f;
        for (size_t i = 0; i < n; ++i) {
            const float a = (float)s[4 * i + 3] * inv;
            px[i] = {{(float)s[4 * i] * inv * a, (float)s[4 * i + 1] * inv * a, (float)s[4 * i + 2] * inv * a, a}};
        }
        return true;
    }
    if (fmt == ND_IMAGE_RGBA32F) {
        const float* s = (const float*)data;
        float lo = INFINITY, hi = -INFINITY;
        for (size_t i = 0; i < n; ++i)
            for (int ch = 0; ch < 3; ++ch) {
                const float v = s[4 * i + ch];
                if (!std::isfinite(v)) continue;
                lo = std::min(lo, v);
                hi = std::max(hi, v);
            }
        const float span = hi > lo ? hi - lo : 1.f;
        for (size_t i = 0; i < n; ++i) {
            Px c = {{0.f, 0.f, 0.f, 1.f}};
            for (int ch = 0; ch < 3; ++ch) {
                const float v = s[4 * i + ch];
                c.c[ch] = std::isfinite(v) ? std::min(std::max((v - lo) / span, 0.f), 1.f) : 0.f;
            }
            px[i] = c;
        }
        return true;
    }
    return false;
}

}  // namespace

extern "C" nd_status nd_image_pack(nd_image_format fmt, const void* pixels, uint32_t w, uint32_t h, uint32_t* words,
                                   size_t words_cap, nd_image_desc* desc) {
    if (!pixels || w == 0 || h == 0 || !desc) return ND_E_INVALID;
    if ((uint64_t)w * h > (1ull << 28)) return ND_E_INVALID;
    std::vector<Px> px;
    if (!to_pixels(fmt, pixels, w, h, px)) return ND_E_INVALID;
    uint32_t factor = 1;
    while (w / factor > kMaxEdge || h / factor > kMaxEdge) ++factor;
    const uint32_t dw = std::max(w / factor, 1u), dh = std::max(h / factor, 1u);
    const size_t n = (size_t)dw * dh;
    desc->width = dw;
    desc->height = dh;
    desc->words = (uint32_t)(4 * n);
    if (!words) return ND_OK;  // size query
    if (words_cap < 4 * n) return ND_E_INVALID;

    std::vector<Px> values(n);
    const float inv_samples = 1.f / (float)(factor * factor);
    for (uint32_t y = 0; y < dh; ++y)
        for (uint32_t x = 0; x < dw; ++x) {
            Px acc = {{0.f, 0.f, 0.f, 0.f}};
            for (uint32_t sy = 0; sy < factor; ++sy) {
                const size_t row = (size_t)std::min(y * factor + sy, h - 1) * w;
                for (uint32_t sx = 0; sx < factor; ++sx) {
                    const Px& p = px[row + std::min(x * factor + sx, w - 1)];
                    for (int ch = 0; ch < 4; ++ch) acc.c[ch] += p.c[ch];
                }
            }
            for (int ch = 0; ch < 4; ++ch) values[(size_t)y * dw + x].c[ch] = q16(acc.c[ch] * inv_samples);
        }

    std::vector<Px> slopes(n);
    float global[4] = {0.f, 0.f, 0.f, 0.f};
    for (uint32_t y = 0; y < dh; ++y)
        for (uint32_t x = 0; x < dw; ++x) {
            const uint32_t x1 = std::min(x + 1, dw - 1), y1 = std::min(y + 1, dh - 1);
            const Px& a = values[(size_t)y * dw + x];
            const Px& b = values[(size_t)y * dw + x1];
            const Px& c = values[(size_t)y1 * dw + x];
            const Px& d = values[(size_t)y1 * dw + x1];
            Px& g = slopes[(size_t)y * dw + x];
            for (int ch = 0; ch < 4; ++ch) {
                const float dx = std::max(std::fabs(b.c[ch] - a.c[ch]), std::fabs(d.c[ch] - c.c[ch]));
                const float dy = std::max(std::fabs(c.c[ch] - a.c[ch]), std::fabs(d.c[ch] - b.c[ch]));
                g.c[ch] = std::sqrt(dx * dx + dy * dy);
                global[ch] = std::max(global[ch], g.c[ch]);
            }
        }
    std::vector<Px> near = slopes;
    dilate(near, dw, dh, kRadiusNear + kDilateMargin);
    dilate(slopes, dw, dh, kRadiusFar + kDilateMargin);
    for (size_t i = 0; i < n; ++i) {
        uint32_t q[4];
        for (int ch = 0; ch < 4; ++ch) q[ch] = (uint32_t)std::round(values[i].c[ch] * 65535.f);
        words[4 * i] = q[0] | (q[1] << 16);
        words[4 * i + 1] = q[2] | (q[3] << 16);
        words[4 * i + 2] = pack_slopes(near[i]);
        words[4 * i + 3] = pack_slopes(slopes[i]);
    }
    for (int ch = 0; ch < 4; ++ch) desc->slope[ch] = global[ch] * kSlopeRoundUp;
    return ND_OK;
}

// Rebuilds the context's device image buffer from its bound slots: headers, then each
// slot's texels.
static nd_status upload_images(nd_ctx* c) {
    size_t words = (size_t)ND_IMAGE_SLOTS * 8;
    for (const auto& im : c->images) words += im.words.size();
    std::vector<uint32_t> all(words, 0u);
    size_t at = (size_t)ND_IMAGE_SLOTS * 8;
    for (int s = 0; s < ND_IMAGE_SLOTS; ++s) {
        const NdImage& im = c->images[s];
        if (im.words.empty()) continue;
        uint32_t* hdr = &all[(size_t)s * 8];
        hdr[0] = (uint32_t)at;
        hdr[1] = im.desc.width;
        hdr[2] = im.desc.height;
        for (int ch = 0; ch < 4; ++ch) std::memcpy(&hdr[3 + ch], &im.desc.slope[ch], 4);
        std::memcpy(&all[at], im.words.data(), 4 * im.words.size());
        at += im.words.size();
    }
    nd_status st = c->image_buf.ensure(c, 4 * words);
    if (st != ND_OK) return st;
    hipError_t e = hipMemcpy(c->image_buf.p, all.data(), 4 * words, hipMemcpyHostToDevice);
    if (e != hipSuccess) return nd_hip_fail(c, e, "upload images");
    c->image_words = all;  // host copy of the device layout: the headers the scene packer reads
    return ND_OK;
}

extern "C" nd_status nd_bind_image(nd_ctx* c, uint32_t slot, nd_image_format fmt, const void* pixels, uint32_t w,
                                   uint32_t h) {
    if (!c) return ND_E_INVALID;
    if (slot >= ND_IMAGE_SLOTS) return nd_fail(c, ND_E_INVALID, "image slot %u out of range (%d slots)", slot, ND_IMAGE_SLOTS);
    if (hipSetDevice(c->device) != hipSuccess) return nd_fail(c, ND_E_DEVICE, "hipSetDevice(%d) failed", c->device);
    NdImage& im = c->images[slot];
    if (!pixels || w == 0 || h == 0) {  // unbind
        im.words.clear();
        im.desc = nd_image_desc{};
        return upload_images(c);
    }
    nd_image_desc d{};
    nd_status st = nd_image_pack(fmt, pixels, w, h, nullptr, 0, &d);
    if (st != ND_OK) return nd_fail(c, st, "image slot %u: format %d, %ux%u is not a valid image", slot, (int)fmt, w, h);
    std::vector<uint32_t> words(d.words);
    if ((st = nd_image_pack(fmt, pixels, w, h, words.data(), words.size(), &d)) != ND_OK)
        return nd_fail(c, st, "image slot %u: packing failed", slot);
    im.words.swap(words);
    im.desc = d;
    return upload_images(c);
}

// The image fields a heightmap formula slot reads (DevInst / DevMixin h[0..8]): texel
// offset, width, height, per-channel slope bound, and the device buffer's address.
void nd_image_hoist(const nd_ctx* c, int slot, float* h) {
    uint32_t hdr[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (slot >= 0 && slot < ND_IMAGE_SLOTS && c->image_words.size() >= (size_t)ND_IMAGE_SLOTS * 8)
        std::memcpy(hdr, &c->image_words[(size_t)slot * 8], sizeof(hdr));
    const uint64_t ptr = (uint64_t)(uintptr_t)c->image_buf.p;
    const uint32_t lo = (uint32_t)ptr, hi = (uint32_t)(ptr >> 32);
    std::memcpy(&h[0], &hdr[0], 4);  // offset
    std::memcpy(&h[1], &hdr[1], 4);  // width (0: no image bound)
    std::memcpy(&h[2], &hdr[2], 4);  // height
    for (int ch = 0; ch < 4; ++ch) std::memcpy(&h[3 + ch], &hdr[3 + ch], 4);
    std::memcpy(&h[7], &lo, 4);
    std::memcpy(&h[8], &hi, 4);
}
