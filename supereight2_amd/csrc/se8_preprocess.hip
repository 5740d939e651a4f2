// Input side of the per-frame loop: se::preprocessor::downsample_depth
// (src/map/preprocessor.cpp:19-84), the app's step between reading a frame and tracking it
// (app/src/main.cpp:118-120), fused with the readers' conversion of stored uint16 depth to metres
// (app/src/reader_raw.cpp:116-121: millimetres / 1000; app/src/reader_tum.cpp, nextDepth:
// value * inverse scale).
//
// One thread per output pixel.  The reference sorts the region's valid depths (>= 1e-5; NaN fails
// the test) and takes region[n / 2]; here the selected element is the one with exactly n / 2
// elements before it in the order (depth, then the region's row-major position), found by counting
// over the region, held in registers for ratios 1, 2 and 4 and read from memory (cached)
// otherwise.  Equal depths are ordered by position; the reference's std::sort leaves their order
// unspecified, which only affects which of several equal-depth pixels the index map names, never
// the depth.
#include "se8_internal.hpp"

using namespace se8;

namespace {

template <typename In>
__device__ __forceinline__ float to_depth(In v, float scale, int divide);

template <>
__device__ __forceinline__ float to_depth<float>(float v, float, int)
{
    return v;
}

template <>
__device__ __forceinline__ float to_depth<uint16_t>(uint16_t v, float scale, int divide)
{
    const float f = (float) v;
    return divide ? f / scale : f * scale;
}

// The reference keeps positive, non-NaN depths.
__device__ __forceinline__ bool valid_depth(float d)
{
    return d >= 1e-5f;
}

// Ratio R known at compile time: the R x R region in registers.
template <int R, typename In>
__global__ __launch_bounds__(256) void k_downsample_fixed(const In* in, int w_in, int w_out, int h_out, float scale,
                                                          int divide, float* out, uint64_t* map)
{
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= w_out || y >= h_out)
        return;
    const int xb = x * R, yb = y * R;
    float d[R * R];
#pragma unroll
    for (int j = 0; j < R; j++)
#pragma unroll
        for (int i = 0; i < R; i++)
            d[j * R + i] = to_depth<In>(in[(size_t) (yb + j) * w_in + xb + i], scale, divide);
    int n = 0;
#pragma unroll
    for (int k = 0; k < R * R; k++)
        n += valid_depth(d[k]) ? 1 : 0;
    float best = 0.f;
    int best_k = 0;
    if (n > 0) {
        const int want = n / 2;
#pragma unroll
        for (int k = 0; k < R * R; k++) {
            int before = 0;
#pragma unroll
            for (int l = 0; l < R * R; l++)
                before += (valid_depth(d[l]) && (d[l] < d[k] || (d[l] == d[k] && l < k))) ? 1 : 0;
            if (valid_depth(d[k]) && before == want) {
                best = d[k];
                best_k = k;
            }
        }
    }
    out[(size_t) y * w_out + x] = best;
    if (map)
        map[(size_t) y * w_out + x] = (uint64_t) (xb + best_k % R) + (uint64_t) (yb + best_k / R) * (uint64_t) w_in;
}

// Any ratio: the region is read from memory (cached) for every candidate.
template <typename In>
__global__ __launch_bounds__(256) void k_downsample_any(const In* in, int w_in, int w_out, int h_out, int R,
                                                        float scale, int divide, float* out, uint64_t* map)
{
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= w_out || y >= h_out)
        return;
    const int xb = x * R, yb = y * R;
    auto at = [&](int k) { return to_depth<In>(in[(size_t) (yb + k / R) * w_in + xb + k % R], scale, divide); };
    int n = 0;
    for (int k = 0; k < R * R; k++)
        n += valid_depth(at(k)) ? 1 : 0;
    float best = 0.f;
    int best_k = 0;
    if (n > 0) {
        const int want = n / 2;
        for (int k = 0; k < R * R; k++) {
            const float dk = at(k);
            if (!valid_depth(dk))
                continue;
            int before = 0;
            for (int l = 0; l < R * R; l++) {
                const float dl = at(l);
                before += (valid_depth(dl) && (dl < dk || (dl == dk && l < k))) ? 1 : 0;
            }
            if (before == want) {
                best = dk;
                best_k = k;
                break;
            }
        }
    }
    out[(size_t) y * w_out + x] = best;
    if (map)
        map[(size_t) y * w_out + x] = (uint64_t) (xb + best_k % R) + (uint64_t) (yb + best_k / R) * (uint64_t) w_in;
}

template <typename In>
int launch_downsample(const In* in, int w_in, int h_in, float scale, int divide, float* out, int w_out, int h_out,
                      uint64_t* map, hipStream_t s)
{
    SE8_REQUIRE(in && out && w_out > 0 && h_out > 0 && w_in >= w_out && h_in >= h_out,
                "downsample_depth: bad image sizes or null buffers");
    SE8_REQUIRE(w_in % w_out == 0 && h_in % h_out == 0 && w_in / w_out == h_in / h_out,
                "downsample_depth: the input must be the same integer multiple of the output in both directions");
    const int R = w_in / w_out;
    const dim3 grid((unsigned) ((w_out + 15) / 16), (unsigned) ((h_out + 15) / 16));
    switch (R) {
    case 1:
        k_downsample_fixed<1, In><<<grid, 256, 0, s>>>(in, w_in, w_out, h_out, scale, divide, out, map);
        break;
    case 2:
        k_downsample_fixed<2, In><<<grid, 256, 0, s>>>(in, w_in, w_out, h_out, scale, divide, out, map);
        break;
    case 4:
        k_downsample_fixed<4, In><<<grid, 256, 0, s>>>(in, w_in, w_out, h_out, scale, divide, out, map);
        break;
    default:
        k_downsample_any<In><<<grid, 256, 0, s>>>(in, w_in, w_out, h_out, R, scale, divide, out, map);
        break;
    }
    SE8_HIP_CHECK(hipGetLastError());
    return SE8_OK;
}

} // namespace

extern "C" int se8_downsample_depth(const float* in_dev, int w_in, int h_in, float* out_dev, int w_out, int h_out,
                                    uint64_t* map_dev, se8_stream stream)
{
    return launch_downsample<float>(in_dev, w_in, h_in, 1.f, 0, out_dev, w_out, h_out, map_dev, (hipStream_t) stream);
}

extern "C" int se8_downsample_depth_u16(const uint16_t* in_dev, int w_in, int h_in, float scale, int divide,
                                        float* out_dev, int w_out, int h_out, uint64_t* map_dev, se8_stream stream)
{
    SE8_REQUIRE(scale > 0.f, "se8_downsample_depth_u16: the scale must be positive");
    return launch_downsample<uint16_t>(in_dev, w_in, h_in, scale, divide ? 1 : 0, out_dev, w_out, h_out, map_dev,
                                       (hipStream_t) stream);
}
