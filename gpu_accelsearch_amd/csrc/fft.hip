// fft.hip -- row transforms, transposes, four-step complex FFT and R2C.
//
// Rows of length 64..16384 run as one workgroup each (fft_block.hpp).  Longer
// power-of-two transforms use the four-step method m = m1*m2: transpose, m2
// row transforms of length m1 with the exp(dir 2 pi i n2 k1 / m) factor fused
// into their store, transpose, m1 row transforms of length m2, transpose.
// A real series of n floats is read as n/2 complex values (even/odd pairs),
// transformed, and split into the n/2+1 bins of the real transform.
#include <cmath>
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

template <int N, int T = FftCfg<N>::T>
static std::vector<cplx> host_twiddles()
{
    using F = BlockFFT<N, T, 1>;
    std::vector<cplx> tw(F::TW_COUNT > 0 ? F::TW_COUNT : 1, mk(1.f, 0.f));
    F::build_twiddles(tw.data());
    return tw;
}

static std::vector<cplx> host_twiddles_for(int n, int t)
{
    if (t != fft_threads(n)) {
        // the other shapes the correlation and segment kernels instantiate
        if (n == 128 && t == 64) return host_twiddles<128, 64>();
        if (n == 1024 && t == 128) return host_twiddles<1024, 128>();
        if (n == 1024 && t == 256) return host_twiddles<1024, 256>();
        if (n == 512 && t == 128) return host_twiddles<512, 128>();
        if (n == 2048 && t == 512) return host_twiddles<2048, 512>();
        if (n == 2048 && t == 256) return host_twiddles<2048, 256>();
        if (n == 4096 && t == 512) return host_twiddles<4096, 512>();
        if (n == 8192 && t == 512) return host_twiddles<8192, 512>();
        if (n == 8192 && t == 1024) return host_twiddles<8192, 1024>();
        if (n == 16384 && t == 1024) return host_twiddles<16384, 1024>();
        return {};
    }
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
const cplx* fft_twiddles(int n, int* err) { return fft_twiddles_t(n, fft_threads(n), err); }

const cplx* fft_twiddles_t(int n, int t, int* err)
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
    auto key = std::make_pair(dev, n * 4096 + t);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    std::vector<cplx> h = host_twiddles_for(n, t);
    if (h.empty()) {
        set_error("no row FFT of length %d with %d threads", n, t);
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

template <int N, int T>
static std::vector<cplx> host_wave_twiddles()
{
    using F = WaveFFT<N, T, 1>;
    std::vector<cplx> tw(F::TW_COUNT);
    F::build_twiddles(tw.data());
    return tw;
}

const cplx* fft_wave_twiddles(int n, int t, int* err)
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
    auto key = std::make_pair(dev, n * 4096 + t);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    std::vector<cplx> h;
    if (n == 8192 && t == 512) h = host_wave_twiddles<8192, 512>();
    else if (n == 4096 && t == 512) h = host_wave_twiddles<4096, 512>();
    else if (n == 4096 && t == 256) h = host_wave_twiddles<4096, 256>();
    else if (n == 2048 && t == 128) h = host_wave_twiddles<2048, 128>();
    if (h.empty()) {
        set_error("no wavefront-split transform of length %d with %d threads", n, t);
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

// Rows of length L into the wavefront-decimated order: out[row][(n mod W) L/W + n/W]
// = in[row][n], W = 2^lw (templates of the wavefront-split correlation).
__global__ __launch_bounds__(256) void decimate_rows_kernel(const cplx* __restrict__ in,
                                                            cplx* __restrict__ out, int L, int lw)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= L) return;
    const size_t row = (size_t)blockIdx.y * L;
    const int W = 1 << lw;
    out[row + (size_t)(n & (W - 1)) * (L >> lw) + (n >> lw)] = in[row + n];
}

int decimate_rows(const cplx* in, cplx* out, int L, int lw, long nrows, hipStream_t s)
{
    if (nrows <= 0 || lw <= 0) return FDAS_OK;
    if (nrows > 65535) {
        set_error("decimate_rows: %ld rows exceed the grid limit", nrows);
        return FDAS_ERR_INVALID;
    }
    hipLaunchKernelGGL(decimate_rows_kernel, dim3((L + 255) / 256, (unsigned)nrows), dim3(256), 0, s,
                       in, out, L, lw);
    return check_launch("decimate_rows_kernel");
}

const cplx* fft_odd_twiddles(int r, int m, int* err)
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
    auto key = std::make_pair(dev, r * 65536 + m);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    if ((r != 3 && r != 5 && r != 7) || m < 64 || m > 4096 || (m & (m - 1))) {
        set_error("no length-%d*%d transform", r, m);
        *err = FDAS_ERR_UNSUPPORTED;
        return nullptr;
    }
    std::vector<cplx> h((size_t)(r - 1) * m);
    build_odd_twiddles(r, m, h.data());
    cplx* d = nullptr;
    hipError_t e = hipMalloc(&d, h.size() * sizeof(cplx));
    if (e == hipSuccess) e = hipMemcpy(d, h.data(), h.size() * sizeof(cplx), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        *err = hip_fail(e, "radix twiddle table", __FILE__, __LINE__);
        return nullptr;
    }
    cache[key] = d;
    return d;
}

int phase_table(long long M, PhaseTable* out)
{
    static std::mutex mu;
    static std::map<std::pair<int, long long>, PhaseTable> cache;
    if (M < 1) {
        set_error("phase table of modulus %lld", M);
        return FDAS_ERR_INVALID;
    }
    int dev = 0;
    FDAS_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    auto key = std::make_pair(dev, M);
    auto it = cache.find(key);
    if (it != cache.end()) {
        *out = it->second;
        return FDAS_OK;
    }
    int shift = 0;
    while ((1LL << (2 * shift)) < M) ++shift;   // 2^shift >= sqrt(M)
    const long long nlo = 1LL << shift, nhi = (M + nlo - 1) / nlo;
    std::vector<cplx> h((size_t)(nlo + nhi));
    const double two_pi = 6.283185307179586476925286766559;
    for (long long i = 0; i < nlo; ++i) {
        const double a = two_pi * (double)i / (double)M;
        h[(size_t)i] = mk((float)std::cos(a), (float)std::sin(a));
    }
    for (long long i = 0; i < nhi; ++i) {
        const double a = two_pi * (double)((i * nlo) % M) / (double)M;
        h[(size_t)(nlo + i)] = mk((float)std::cos(a), (float)std::sin(a));
    }
    cplx* d = nullptr;
    FDAS_HIP(hipMalloc(&d, h.size() * sizeof(cplx)));
    FDAS_HIP(hipMemcpy(d, h.data(), h.size() * sizeof(cplx), hipMemcpyHostToDevice));
    PhaseTable t;
    t.lo = d;
    t.hi = d + nlo;
    t.shift = shift;
    cache[key] = t;
    *out = t;
    return FDAS_OK;
}

// ------------------------------------------------------------------ kernels
// Forward transform of rows of length R*M into the decimated order (template
// banks of the length-R*M correlations).  Every thread's loads precede the first
// barrier of the transform and its stores follow the last, so in place is safe.
template <int R, int M>
constexpr int odd_rows_threads()
{
    return M <= 512 ? 64 : R == 3 && M == 4096 ? 256 : M / 8;
}

template <int R, int M>
__global__ __launch_bounds__((odd_rows_threads<R, M>())) void fft_odd_rows_kernel(
    const cplx* in, cplx* out, const cplx* __restrict__ tw, const cplx* __restrict__ twr)
{
    constexpr int T = odd_rows_threads<R, M>();
    using F = BlockFFT<M, T, -1>;
    constexpr int E = F::E;
    static_assert(F::NPASS >= 2, "in-place rows need a barrier between loads and stores");
    __shared__ cplx lds[R * F::LDS_ELEMS];
    const int t = threadIdx.x;
    const cplx* x = in + (size_t)blockIdx.x * R * M;
    cplx* y = out + (size_t)blockIdx.x * R * M;
    cplx v[R][E];
#pragma unroll
    for (int a = 0; a < R; ++a)
#pragma unroll
        for (int i = 0; i < E; ++i) v[a][i] = x[F::in_index(t, i) + M * a];
    RadixOdd<R, M, T>::fwd_split(v, t, twr);
    F::template run_multi<R>(v, lds, tw, t);
#pragma unroll
    for (int a = 0; a < R; ++a)
#pragma unroll
        for (int i = 0; i < E; ++i) y[M * a + F::out_index(t, i)] = v[a][i];
}

template <int R, int M>
static int fft_odd_rows_t(const cplx* in, cplx* out, long nrows, hipStream_t s)
{
    constexpr int T = odd_rows_threads<R, M>();
    int err = FDAS_OK;
    const cplx* tw = fft_twiddles_t(M, T, &err);
    if (!tw) return err;
    const cplx* twr = fft_odd_twiddles(R, M, &err);
    if (!twr) return err;
    hipLaunchKernelGGL((fft_odd_rows_kernel<R, M>), dim3((unsigned)nrows), dim3(T), 0, s, in, out, tw,
                       twr);
    return check_launch("fft_odd_rows_kernel");
}

int fft_odd_rows(const cplx* in, cplx* out, int r, int m, long nrows, hipStream_t s)
{
    if (nrows <= 0) return FDAS_OK;
    if (nrows > 2147483647L) {
        set_error("fft_odd_rows: %ld rows exceed the grid limit", nrows);
        return FDAS_ERR_INVALID;
    }
    switch (r * 65536 + m) {
        case 3 * 65536 + 256: return fft_odd_rows_t<3, 256>(in, out, nrows, s);
        case 3 * 65536 + 512: return fft_odd_rows_t<3, 512>(in, out, nrows, s);
        case 3 * 65536 + 1024: return fft_odd_rows_t<3, 1024>(in, out, nrows, s);
        case 3 * 65536 + 2048: return fft_odd_rows_t<3, 2048>(in, out, nrows, s);
        case 3 * 65536 + 4096: return fft_odd_rows_t<3, 4096>(in, out, nrows, s);
        case 5 * 65536 + 256: return fft_odd_rows_t<5, 256>(in, out, nrows, s);
        case 5 * 65536 + 512: return fft_odd_rows_t<5, 512>(in, out, nrows, s);
        case 5 * 65536 + 1024: return fft_odd_rows_t<5, 1024>(in, out, nrows, s);
        case 5 * 65536 + 2048: return fft_odd_rows_t<5, 2048>(in, out, nrows, s);
        case 7 * 65536 + 1024: return fft_odd_rows_t<7, 1024>(in, out, nrows, s);
        default:
            set_error("no length-%d*%d row transform", r, m);
            return FDAS_ERR_UNSUPPORTED;
    }
}

template <int N, int DIR>
__global__ __launch_bounds__(FftCfg<N>::T) void fft_rows_kernel(
    const cplx* __restrict__ in, cplx* __restrict__ out, long in_stride, long out_stride,
    const cplx* __restrict__ tw, float scale, long nrows, long tw_m, PhaseTable ph)
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
            cplx w = phase_at(ph, ((long long)row * k) % tw_m);
            if (DIR < 0) w.y = -w.y;
            y = cmul(y, w);
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
                                                        cplx* __restrict__ X, long m, PhaseTable ph)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k > m) return;
    const cplx a = Z[k == m ? 0 : k];
    const cplx bc = Z[k == 0 ? 0 : m - k];
    const cplx b = mk(bc.x, -bc.y);
    const cplx fe = mk(0.5f * (a.x + b.x), 0.5f * (a.y + b.y));
    const cplx d = mk(a.x - b.x, a.y - b.y);
    const cplx fo = mk(0.5f * d.y, -0.5f * d.x);
    cplx w = phase_at(ph, k);   // exp(-pi i k / m): the table has modulus 2m
    w.y = -w.y;
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
    PhaseTable ph;
    if (tw_m > 0) FDAS_TRY(phase_table(tw_m, &ph));
    dim3 grid((unsigned)nrows), block(FftCfg<N>::T);
    if (dir < 0)
        hipLaunchKernelGGL((fft_rows_kernel<N, -1>), grid, block, 0, s, in, out, in_stride,
                           out_stride, tw, scale, nrows, tw_m, ph);
    else
        hipLaunchKernelGGL((fft_rows_kernel<N, 1>), grid, block, 0, s, in, out, in_stride,
                           out_stride, tw, scale, nrows, tw_m, ph);
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

// ------------------------------------------------------------------ Bluestein
// m-point forward DFT of any length through power-of-two transforms of length
// p >= 2m-1:  Z[k] = conj(c_k) * sum_j (z_j conj(c_j)) c_{k-j},  c_j = exp(i pi j^2/m).
// The chirp phase uses j^2 mod 2m in integers, so it is exact for any m; the
// phase table has modulus 2m.
__device__ __forceinline__ cplx chirp(long long j, long long m, const PhaseTable& ph)
{
    return phase_at(ph, (j * j) % (2 * m));
}

// a[j] = z[j] conj(c_j) for j < m, 0 for m <= j < p
__global__ __launch_bounds__(256) void bluestein_pre_kernel(const cplx* __restrict__ z,
                                                            cplx* __restrict__ a, long m, long p,
                                                            PhaseTable ph)
{
    const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= p) return;
    a[j] = j < m ? cmul(z[j], conjg(chirp(j, m, ph))) : mk(0.f, 0.f);
}

// b[j] = c_j for j < m, b[p-j] = c_j for 0 < j < m, zero elsewhere
__global__ __launch_bounds__(256) void bluestein_kernel_kernel(cplx* __restrict__ b, long m, long p,
                                                              PhaseTable ph)
{
    const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= p) return;
    cplx v = mk(0.f, 0.f);
    if (j < m) v = chirp(j, m, ph);
    else if (j > p - m) v = chirp(p - j, m, ph);
    b[j] = v;
}

__global__ __launch_bounds__(256) void pointwise_mul_kernel(cplx* __restrict__ a,
                                                            const cplx* __restrict__ b, long p,
                                                            float scale)
{
    const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= p) return;
    const cplx v = cmul(a[j], b[j]);
    a[j] = mk(v.x * scale, v.y * scale);
}

__global__ __launch_bounds__(256) void bluestein_post_kernel(const cplx* __restrict__ y,
                                                             cplx* __restrict__ Z, long m,
                                                             PhaseTable ph)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    Z[k] = cmul(y[k], conjg(chirp(k, m, ph)));
}

static unsigned blocks_for(long n) { return (unsigned)((n + 255) / 256); }

int r2c_plan_create(long n, R2CPlan* rp)
{
    *rp = R2CPlan();
    if (n < 256 || (n & 1)) {
        set_error("real transform length %ld must be even and >= 256", n);
        return FDAS_ERR_UNSUPPORTED;
    }
    rp->n = n;
    rp->m = n / 2;
    rp->pow2 = is_pow2(rp->m);
    if (rp->pow2) {
        rp->p = rp->m;
        rp->work_elems = (size_t)3 * rp->m;
        return FDAS_OK;
    }
    rp->p = 1;
    while (rp->p < 2 * rp->m - 1) rp->p <<= 1;
    if (rp->p > (1L << 28)) {
        set_error("real transform length %ld too long for the chirp-z path", n);
        return FDAS_ERR_UNSUPPORTED;
    }
    // Z (m) + a (p) + transformed chirp (p) + four-step scratch (2p)
    rp->work_elems = (size_t)rp->m + 4 * (size_t)rp->p;
    return FDAS_OK;
}

int fft_r2c_plan(const R2CPlan& rp, const float* in, cplx* out, cplx* work, hipStream_t s)
{
    if (((uintptr_t)in) & 7) {
        set_error("real transform input must be 8-byte aligned");
        return FDAS_ERR_INVALID;
    }
    const long m = rp.m;
    cplx* Z = work;
    const cplx* z = reinterpret_cast<const cplx*>(in);
    if (rp.pow2) {
        FDAS_TRY(fft_c2c(z, Z, m, -1, work + m, s));
    } else {
        // chirp-z: a = x w, A = FFT(a) FFT(b) / p, Z = conj(w) IFFT(A); the
        // transformed chirp b is rebuilt per call, so a plan holds no device state
        const long p = rp.p;
        cplx* a = work + m;
        cplx* kb = a + p;
        cplx* scratch = kb + p;
        PhaseTable ch;
        FDAS_TRY(phase_table(2LL * m, &ch));
        hipLaunchKernelGGL(bluestein_kernel_kernel, dim3(blocks_for(p)), dim3(256), 0, s, kb, m, p, ch);
        FDAS_TRY(check_launch("bluestein_kernel_kernel"));
        FDAS_TRY(fft_c2c(kb, kb, p, -1, scratch, s));
        hipLaunchKernelGGL(bluestein_pre_kernel, dim3(blocks_for(p)), dim3(256), 0, s, z, a, m, p, ch);
        FDAS_TRY(check_launch("bluestein_pre_kernel"));
        FDAS_TRY(fft_c2c(a, a, p, -1, scratch, s));
        hipLaunchKernelGGL(pointwise_mul_kernel, dim3(blocks_for(p)), dim3(256), 0, s, a, kb, p,
                           (float)(1.0 / (double)p));
        FDAS_TRY(check_launch("pointwise_mul_kernel"));
        FDAS_TRY(fft_c2c(a, a, p, +1, scratch, s));
        hipLaunchKernelGGL(bluestein_post_kernel, dim3(blocks_for(m)), dim3(256), 0, s, a, Z, m, ch);
        FDAS_TRY(check_launch("bluestein_post_kernel"));
    }
    PhaseTable sp;
    FDAS_TRY(phase_table(2LL * m, &sp));
    hipLaunchKernelGGL(r2c_split_kernel, dim3(blocks_for(m + 1)), dim3(256), 0, s, Z, out, m, sp);
    return check_launch("r2c_split_kernel");
}

}  // namespace fdas
