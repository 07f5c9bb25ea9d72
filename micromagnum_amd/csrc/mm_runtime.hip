// mm_runtime.hip -- library globals of libmmhip.so: error reporting, device
// selection, memory helpers, demag-tensor setup entry points, tableaus.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "mm_conv.h"

static thread_local std::string g_last_error;

void mm_set_error(const std::string &msg) { g_last_error = msg; }

extern "C" const char *mm_last_error(void) { return g_last_error.c_str(); }

extern "C" int mm_init(int device)
{
    MM_CAPI_BEGIN
    int count = 0;
    MM_HIP(hipGetDeviceCount(&count));
    MM_REQUIRE(count > 0, "no HIP device visible");
    MM_REQUIRE(device >= 0 && device < count, "device index out of range");
    MM_HIP(hipSetDevice(device));
    MM_HIP(hipFree(nullptr));
    (void)mm_max_lds_bytes();
    mm_noalloc_setup();
    MM_CAPI_END
}

extern "C" int mm_device_alloc(size_t bytes, void **out)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(out, "null output");
    *out = nullptr;
    if (bytes) MM_HIP(hipMalloc(out, bytes));
    MM_CAPI_END
}

extern "C" int mm_device_free(void *ptr)
{
    MM_CAPI_BEGIN
    if (ptr) MM_HIP(hipFree(ptr));
    MM_CAPI_END
}

extern "C" int mm_memcpy_htod(void *dst, const void *src_host, size_t bytes, void *stream)
{
    MM_CAPI_BEGIN
    MM_HIP(hipMemcpyAsync(dst, src_host, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
    MM_HIP(hipStreamSynchronize((hipStream_t)stream));
    MM_CAPI_END
}

extern "C" int mm_memcpy_dtoh(void *dst_host, const void *src, size_t bytes, void *stream)
{
    MM_CAPI_BEGIN
    MM_HIP(hipMemcpyAsync(dst_host, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
    MM_HIP(hipStreamSynchronize((hipStream_t)stream));
    MM_CAPI_END
}

// plain streaming copy (16-byte vectors); also the reference kernel for
// converting hardware byte counters into bytes of a known transfer
__global__ void k_copy_calib(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n16)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

extern "C" int mm_device_copy(void *dst, const void *src, size_t bytes, void *stream)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(dst && src && bytes % 16 == 0, "copy needs 16-byte multiples");
    const size_t n16 = bytes / 16;
    const int grid = (int)std::min<size_t>((n16 + 255) / 256, 8192);
    hipLaunchKernelGGL(k_copy_calib, dim3(grid > 0 ? grid : 1), dim3(256), 0, (hipStream_t)stream,
                       (const uint4 *)src, (uint4 *)dst, n16);
    MM_HIP(hipGetLastError());
    MM_CAPI_END
}

extern "C" int mm_synchronize(void *stream)
{
    MM_CAPI_BEGIN
    MM_HIP(hipStreamSynchronize((hipStream_t)stream));
    MM_CAPI_END
}

extern "C" int mm_round_tensor_dimension(int dim, int periodic, int padding, int *out)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(out && dim >= 1, "bad argument");
    *out = mm_round_tensor_dimension_impl(dim, periodic != 0, padding);
    MM_CAPI_END
}

// Cache file of a tensor field, named as the reference names it
// (demag_tensor.cpp:52-71) so that caches can be shared: raw doubles in the
// returned layout, cell sizes in integer picometres.
static std::string tensor_cache_file(const char *dir, int nx, int ny, int nz, int ex, int ey, int ez, double dx,
                                     double dy, double dz, int px, int py, int pz, int repeat)
{
    std::ostringstream ss;
    ss << dir << "/Demag--" << nx << "-" << ny << "-" << nz << "--" << ex << "-" << ey << "-" << ez << "--"
       << int(dx * 1e12) << "-" << int(dy * 1e12) << "-" << int(dz * 1e12);
    if (px || py || pz)
        ss << "--p-" << (px ? "x" : "") << (py ? "y" : "") << (pz ? "z" : "") << "-" << repeat;
    ss << ".dat";
    return ss.str();
}

extern "C" int mm_generate_demag_tensor(int nx, int ny, int nz, double dx, double dy, double dz, int px, int py,
                                        int pz, int repeat, int padding, const char *cache_dir, int nthreads,
                                        double *N_host, int *exp_out)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1, "bad mesh size");
    MM_REQUIRE(dx > 0 && dy > 0 && dz > 0, "cell sizes must be positive");
    MM_REQUIRE(!(px || py || pz) || repeat >= 1, "periodic_repeat must be >= 1");
    const int ex = mm_round_tensor_dimension_impl(nx, px != 0, padding);
    const int ey = mm_round_tensor_dimension_impl(ny, py != 0, padding);
    const int ez = mm_round_tensor_dimension_impl(nz, pz != 0, padding);
    if (exp_out) {
        exp_out[0] = ex;
        exp_out[1] = ey;
        exp_out[2] = ez;
    }
    if (!N_host) return MM_OK;
    const size_t count = (size_t)6 * ex * ey * ez;
    // demag_tensor.cpp:146-152: skip the computation (contents undefined there, zeros here)
    if (std::getenv("MAGNUM_DEMAG_GARBAGE")) {
        std::fill(N_host, N_host + count, 0.0);
        return MM_OK;
    }
    std::string path;
    if (cache_dir && cache_dir[0]) {
        path = tensor_cache_file(cache_dir, nx, ny, nz, ex, ey, ez, dx, dy, dz, px, py, pz, repeat);
        std::ifstream in(path, std::ios::binary);
        if (in) {
            in.read(reinterpret_cast<char *>(N_host), (std::streamsize)(sizeof(double) * count));
            if ((size_t)in.gcount() == sizeof(double) * count) return MM_OK;
        }
    }
    const auto t0 = std::chrono::steady_clock::now();
    mm_demag_tensor_impl(nx, ny, nz, dx, dy, dz, ex, ey, ez, N_host, nthreads, px, py, pz, repeat);
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    // saved when the computation took longer than 30 s (demag_tensor.cpp:179-182),
    // or always with MM_DEMAG_CACHE_ALWAYS set
    if (!path.empty() && (secs > 30.0 || std::getenv("MM_DEMAG_CACHE_ALWAYS"))) {
        const std::string tmp = path + ".tmp";
        std::ofstream out(tmp, std::ios::binary);
        out.write(reinterpret_cast<const char *>(N_host), (std::streamsize)(sizeof(double) * count));
        out.close();
        if (out) std::rename(tmp.c_str(), path.c_str());
        else std::remove(tmp.c_str());
    }
    MM_CAPI_END
}

extern "C" int mm_demag_tensor_entry(int i, int j, int k, double dx, double dy, double dz, double *v6)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(i >= 0 && j >= 0 && k >= 0 && v6, "bad argument");
    MM_REQUIRE(dx > 0 && dy > 0 && dz > 0, "cell sizes must be positive");
    mm_demag_tensor_entry_impl(i, j, k, dx, dy, dz, v6);
    MM_CAPI_END
}

// tableaus.py:54-139; classical RK4 for MM_RK4; euler.py:27-36 for MM_EULER
extern "C" int mm_tableau(int method, mm_tableau_t *t)
{
    MM_CAPI_BEGIN
    MM_REQUIRE(t, "null output");
    std::memset(t, 0, sizeof(*t));
    switch (method) {
    case MM_RKF45: {
        t->num_steps = 6; t->order = 4; t->fsal = 0;
        const double a[6] = {0.0, 1.0 / 4.0, 3.0 / 8.0, 12.0 / 13.0, 1.0, 1.0 / 2.0};
        const double b[6][5] = {{0}, {1.0 / 4.0}, {3.0 / 32.0, 9.0 / 32.0},
                                {1932.0 / 2197.0, -7200.0 / 2197.0, 7296.0 / 2197.0},
                                {439.0 / 216.0, -8.0, 3680.0 / 513.0, -845.0 / 4104.0},
                                {-8.0 / 27.0, 2.0, -3544.0 / 2565.0, 1859.0 / 4104.0, -11.0 / 40.0}};
        const double c4[6] = {25.0 / 216.0, 0.0, 1408.0 / 2565.0, 2197.0 / 4104.0, -1.0 / 5.0, 0.0};
        const double c5[6] = {16.0 / 135.0, 0.0, 6656.0 / 12825.0, 28561.0 / 56430.0, -9.0 / 50.0, 2.0 / 55.0};
        for (int i = 0; i < 6; ++i) {
            t->a[i] = a[i]; t->c[i] = c4[i]; t->ec[i] = c4[i] - c5[i];
            for (int j = 0; j < 5; ++j) t->b[i][j] = b[i][j];
        }
        break;
    }
    case MM_CC45: {
        t->num_steps = 6; t->order = 4; t->fsal = 0;
        const double a[6] = {0.0, 1.0 / 5.0, 3.0 / 10.0, 3.0 / 5.0, 1.0, 7.0 / 8.0};
        const double b[6][5] = {{0}, {1.0 / 5.0}, {3.0 / 40.0, 9.0 / 40.0}, {3.0 / 10.0, -9.0 / 10.0, 6.0 / 5.0},
                                {-11.0 / 54.0, 5.0 / 2.0, -70.0 / 27.0, 35.0 / 27.0},
                                {1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0, 44275.0 / 110592.0, 253.0 / 4096.0}};
        const double c4[6] = {2825.0 / 27648.0, 0.0, 18575.0 / 48384.0, 13525.0 / 55296.0, 277.0 / 14336.0, 1.0 / 4.0};
        const double c5[6] = {37.0 / 378.0, 0.0, 250.0 / 621.0, 125.0 / 594.0, 0.0, 512.0 / 1771.0};
        for (int i = 0; i < 6; ++i) {
            t->a[i] = a[i]; t->c[i] = c4[i]; t->ec[i] = c4[i] - c5[i];
            for (int j = 0; j < 5; ++j) t->b[i][j] = b[i][j];
        }
        break;
    }
    case MM_DP54: {
        t->num_steps = 7; t->order = 5; t->fsal = 1;
        const double a[7] = {0.0, 1.0 / 5.0, 3.0 / 10.0, 4.0 / 5.0, 8.0 / 9.0, 1.0, 1.0};
        const double b[7][6] = {{0}, {1.0 / 5.0}, {3.0 / 40.0, 9.0 / 40.0}, {44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0},
                                {19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0},
                                {9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0},
                                {35.0 / 384.0, 0.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0}};
        const double c4[7] = {5179.0 / 57600.0, 0.0, 7571.0 / 16695.0, 393.0 / 640.0, -92097.0 / 339200.0,
                              187.0 / 2100.0, 1.0 / 40.0};
        const double c5[7] = {35.0 / 384.0, 0.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0, 0.0};
        for (int i = 0; i < 7; ++i) {
            t->a[i] = a[i]; t->c[i] = c5[i]; t->ec[i] = c4[i] - c5[i];
            for (int j = 0; j < 6; ++j) t->b[i][j] = b[i][j];
        }
        break;
    }
    case MM_RK23: {
        t->num_steps = 4; t->order = 3; t->fsal = 1;
        const double a[4] = {0.0, 0.5, 3.0 / 4.0, 1.0};
        const double b[4][3] = {{0}, {1.0 / 2.0}, {0.0, 3.0 / 4.0}, {2.0 / 9.0, 1.0 / 3.0, 4.0 / 9.0}};
        const double c2[4] = {7.0 / 24.0, 1.0 / 4.0, 1.0 / 3.0, 1.0 / 8.0};
        const double c3[4] = {2.0 / 9.0, 1.0 / 3.0, 4.0 / 9.0, 0.0};
        for (int i = 0; i < 4; ++i) {
            t->a[i] = a[i]; t->c[i] = c3[i]; t->ec[i] = c2[i] - c3[i];
            for (int j = 0; j < 3; ++j) t->b[i][j] = b[i][j];
        }
        break;
    }
    case MM_EULER: {  // euler.py:27-36
        t->num_steps = 1; t->order = 1; t->fsal = 0;
        t->c[0] = 1.0;
        break;
    }
    case MM_RK4: {
        t->num_steps = 4; t->order = 4; t->fsal = 0;
        t->a[1] = 0.5; t->a[2] = 0.5; t->a[3] = 1.0;
        t->b[1][0] = 0.5; t->b[2][1] = 0.5; t->b[3][2] = 1.0;
        t->c[0] = 1.0 / 6.0; t->c[1] = 2.0 / 6.0; t->c[2] = 2.0 / 6.0; t->c[3] = 1.0 / 6.0;
        break;
    }
    default:
        throw mm_error(MM_ERR_INVALID, "unknown Runge-Kutta method");
    }
    MM_CAPI_END
}
