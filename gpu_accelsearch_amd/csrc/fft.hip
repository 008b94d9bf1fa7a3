// fft.hip -- row transforms, transposes, four-step complex FFT and R2C.
//
// Rows of length 64..16384 run as one workgroup each (fft_block.hpp).  Longer
// power-of-two transforms use the four-step method m = m1*m2: transpose, m2
// row transforms of length m1 with the exp(dir 2 pi i n2 k1 / m) factor fused
// into their store, transpose, m1 row transforms of length m2, transpose.
// A real series of n floats is read as n/2 complex values (even/odd pairs),
// transformed, and split into the n/2+1 bins of the real transform.
#include <map>
#include <mutex>

#include "fdas_internal.hpp"
#include "fft_block.hpp"

namespace fdas {

template <int N>
struct FftCfg;
template <> struct FftCfg<64>    { static constexpr int T = 16; };
template <> struct FftCfg<128>   { static constexpr int T = 32; };
template <> struct FftCfg<256>   { static constexpr int T = 64; };
template <> struct FftCfg<512>   { static constexpr int T = 64; };
template <> struct FftCfg<1024>  { static constexpr int T = 64; };
template <> struct FftCfg<2048>  { static constexpr int T = 128; };
template <> struct FftCfg<4096>  { static constexpr int T = 256; };
template <> struct FftCfg<8192>  { static constexpr int T = 256; };
template <> struct FftCfg<16384> { static constexpr int T = 512; };

int fft_threads(int n)
{
    switch (n) {
        case 64: return FftCfg<64>::T;
        case 128: return FftCfg<128>::T;
        case 256: return FftCfg<256>::T;
        case 512: return FftCfg<512>::T;
        case 1024: return FftCfg<1024>::T;
        case 2048: return FftCfg<2048>::T;
        case 4096: return FftCfg<4096>::T;
        case 8192: return FftCfg<8192>::T;
        case 16384: return FftCfg<16384>::T;
        default: return 0;
    }
}

bool fft_row_supported(long n) { return fft_threads((int)n) > 0 && n <= 16384; }

template <int N>
static std::vector<cplx> host_twiddles()
{
    using F = BlockFFT<N, FftCfg<N>::T, 1>;
    std::vector<cplx> tw(F::TW_COUNT > 0 ? F::TW_COUNT : 1, mk(1.f, 0.f));
    F::build_twiddles(tw.data());
    return tw;
}

static std::vector<cplx> host_twiddles_for(int n)
{
    switch (n) {
        case 64: return host_twiddles<64>();
        case 128: return host_twiddles<128>();
        case 256: return host_twiddles<256>();
        case 512: return host_twiddles<512>();
        case 1024: return host_twiddles<1024>();
        case 2048: return host_twiddles<2048>();
        case 4096: return host_twiddles<4096>();
        case 8192: return host_twiddles<8192>();
        case 16384: return host_twiddles<16384>();
        default: return {};
    }
}

// Twiddle tables live for the life of the process, one set per device.
const cplx* fft_twiddles(int n, int* err)
{
    static std::mutex mu;
    static std::map<std::pair<int, int>, cplx*> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        set_error("hipGetDevice failed");
        *err = FDAS_ERR_HIP;
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(mu);
    auto key = std::make_pair(dev, n);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    std::vector<cplx> h = host_twiddles_for(n);
    if (h.empty()) {
        set_error("no row FFT of length %d", n);
        *err = FDAS_ERR_UNSUPPORTED;
        return nullptr;
    }
    cplx* d = nullptr;
    hipError_t e = hipMalloc(&d, h.size() * sizeof(cplx));
    if (e == hipSuccess) e = hipMemcpy(d, h.data(), h.size() * sizeof(cplx), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        *err = hip_fail(e, "twiddle table", __FILE__, __LINE__);
        return nullptr;
    }
    cache[key] = d;
    return d;
}

// ------------------------------------------------------------------ kernels
template <int N, int DIR>
__global__ __launch_bounds__(FftCfg<N>::T) void fft_rows_kernel(
    const cplx* __restrict__ in, cplx* __restrict__ out, long in_stride, long out_stride,
    const cplx* __restrict__ tw, float scale, long nrows, long tw_m)
{
    constexpr int T = FftCfg<N>::T;
    using F = BlockFFT<N, T, DIR>;
    __shared__ cplx lds[F::LDS_ELEMS];
    const long row = blockIdx.x;
    if (row >= nrows) return;
    const int t = threadIdx.x;
    const cplx* src = in + row * in_stride;
    cplx v[F::E];
#pragma unroll
    for (int i = 0; i < F::E; ++i) v[i] = src[F::in_index(t, i)];
    F::run(v, lds, tw, t);
    cplx* dst = out + row * out_stride;
#pragma unroll
    for (int i = 0; i < F::E; ++i) {
        const int k = F::out_index(t, i);
        cplx y = scale == 1.0f ? v[i] : mk(v[i].x * scale, v[i].y * scale);
        if (tw_m > 0) {
            const long long m = ((long long)row * k) % tw_m;
            double sn, cs;
            sincospi(2.0 * (double)m / (double)tw_m, &sn, &cs);
            y = cmul(y, mk((float)cs, (float)(DIR * sn)));
        }
        dst[k] = y;
    }
}

__global__ __launch_bounds__(256) void transpose_kernel(const cplx* __restrict__ in,
                                                        cplx* __restrict__ out, long rows,
                                                        long cols)
{
    __shared__ cplx tile[32][33];
    const long bx = (long)blockIdx.x * 32, by = (long)blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int j = ty; j < 32; j += 8) {
        const long r = by + j, c = bx + tx;
        if (r < rows && c < cols) tile[j][tx] = in[r * cols + c];
    }
    __syncthreads();
    for (int j = ty; j < 32; j += 8) {
        const long c = bx + j, r = by + tx;
        if (c < cols && r < rows) out[c * rows + r] = tile[tx][j];
    }
}

// X[k] = (Z[k] + conj Z[m-k])/2 - i/2 exp(-2 pi i k / 2m) (Z[k] - conj Z[m-k]),
// k = 0..m, for Z the m-point transform of z[j] = x[2j] + i x[2j+1].
__global__ __launch_bounds__(256) void r2c_split_kernel(const cplx* __restrict__ Z,
                                                        cplx* __restrict__ X, long m)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k > m) return;
    const cplx a = Z[k == m ? 0 : k];
    const cplx bc = Z[k == 0 ? 0 : m - k];
    const cplx b = mk(bc.x, -bc.y);
    const cplx fe = mk(0.5f * (a.x + b.x), 0.5f * (a.y + b.y));
    const cplx d = mk(a.x - b.x, a.y - b.y);
    const cplx fo = mk(0.5f * d.y, -0.5f * d.x);
    double sn, cs;
    sincospi(-(double)k / (double)m, &sn, &cs);
    const cplx w = mk((float)cs, (float)sn);
    X[k] = fe + cmul(w, fo);
}

// ------------------------------------------------------------------ launchers
template <int N>
static int launch_rows(const cplx* in, cplx* out, long nrows, long in_stride, long out_stride,
                       int dir, float scale, long tw_m, hipStream_t s)
{
    int err = FDAS_OK;
    const cplx* tw = fft_twiddles(N, &err);
    if (!tw) return err;
    if (nrows <= 0) return FDAS_OK;
    dim3 grid((unsigned)nrows), block(FftCfg<N>::T);
    if (dir < 0)
        hipLaunchKernelGGL((fft_rows_kernel<N, -1>), grid, block, 0, s, in, out, in_stride,
                           out_stride, tw, scale, nrows, tw_m);
    else
        hipLaunchKernelGGL((fft_rows_kernel<N, 1>), grid, block, 0, s, in, out, in_stride,
                           out_stride, tw, scale, nrows, tw_m);
    return check_launch("fft_rows_kernel");
}

int fft_rows(const cplx* in, cplx* out, int n, long nrows, long in_stride, long out_stride,
             int dir, float scale, long tw_m, hipStream_t s)
{
    if (nrows > 0x7fffffffL) {
        set_error("fft_rows: %ld rows exceed the grid limit", nrows);
        return FDAS_ERR_UNSUPPORTED;
    }
    switch (n) {
        case 64: return launch_rows<64>(in, out, nrows, in_stride, out_stride, dir, scale, tw_m, s);
        case 128: return launch_rows<128>(in, out, nrows, in_stride, out_stride, dir, scale, tw_m, s);
        case 256: return launch_rows<256>(in, out, nrows, in_stride, out_stride, dir, scale, tw_m, s);
        case 512: return launch_rows<512>(in, out, nrows, in_stride, out_stride, dir, scale, tw_m, s);
        case 1024: return launch_rows<1024>(in, out, nrows, in_stride, out_stride, dir, scale, tw_m, s);
        case 2048: return launch_rows<2048>(in, out, nrows, in_stride, out_stride, dir, scale, tw_m, s);
        case 4096: return launch_rows<4096>(in, out, nrows, in_stride, out_stride, dir, scale, tw_m, s);
        case 8192: return launch_rows<8192>(in, out, nrows, in_stride, out_stride, dir, scale, tw_m, s);
        case 16384: return launch_rows<16384>(in, out, nrows, in_stride, out_stride, dir, scale, tw_m, s);
        default:
            set_error("no row FFT of length %d", n);
            return FDAS_ERR_UNSUPPORTED;
    }
}

int transpose(const cplx* in, cplx* out, long rows, long cols, hipStream_t s)
{
    dim3 grid((unsigned)((cols + 31) / 32), (unsigned)((rows + 31) / 32)), block(256);
    hipLaunchKernelGGL(transpose_kernel, grid, block, 0, s, in, out, rows, cols);
    return check_launch("transpose_kernel");
}

static bool is_pow2(long n) { return n > 0 && (n & (n - 1)) == 0; }

int fft_c2c(const cplx* in, cplx* out, long m, int dir, cplx* work, hipStream_t s)
{
    if (!is_pow2(m) || m < 64 || m > (1L << 28)) {
        set_error("fft_c2c: length %ld is not a power of two in [64, 2^28]", m);
        return FDAS_ERR_UNSUPPORTED;
    }
    if (m <= 16384) return fft_rows(in, out, (int)m, 1, m, m, dir, 1.0f, 0, s);
    const int lg = ilog2((int)(m >> 14)) + 14;   // log2(m)
    const long m1 = 1L << (lg / 2), m2 = m / m1;
    cplx* w1 = work;
    cplx* w2 = work + m;
    FDAS_TRY(transpose(in, w1, m1, m2, s));                                   // [m2][m1]
    FDAS_TRY(fft_rows(w1, w2, (int)m1, m2, m1, m1, dir, 1.0f, m, s));          // T[n2][k1]
    FDAS_TRY(transpose(w2, w1, m2, m1, s));                                   // [m1][m2]
    FDAS_TRY(fft_rows(w1, w2, (int)m2, m1, m2, m2, dir, 1.0f, 0, s));          // [k1][k2]
    FDAS_TRY(transpose(w2, out, m1, m2, s));                                  // [k2][k1]
    return FDAS_OK;
}

size_t fft_r2c_work_elems(long n) { return (size_t)3 * (size_t)(n / 2); }

int fft_r2c(const float* in, cplx* out, long n, cplx* work, hipStream_t s)
{
    if (n < 128 || !is_pow2(n)) {
        set_error("fft_r2c: length %ld is not a power of two >= 128", n);
        return FDAS_ERR_UNSUPPORTED;
    }
    if (((uintptr_t)in) & 7) {
        set_error("fft_r2c: input must be 8-byte aligned");
        return FDAS_ERR_INVALID;
    }
    const long m = n / 2;
    cplx* Z = work;              // m
    cplx* scratch = work + m;    // 2m
    FDAS_TRY(fft_c2c(reinterpret_cast<const cplx*>(in), Z, m, -1, scratch, s));
    dim3 grid((unsigned)((m + 1 + 255) / 256)), block(256);
    hipLaunchKernelGGL(r2c_split_kernel, grid, block, 0, s, Z, out, m);
    return check_launch("r2c_split_kernel");
}

}  // namespace fdas
