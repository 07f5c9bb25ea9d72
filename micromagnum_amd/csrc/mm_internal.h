// mm_internal.h -- declarations shared by the translation units of libmmhip.so.
#ifndef MM_INTERNAL_H
#define MM_INTERNAL_H

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "../../include/mm_hotpath.h"

// error handling: every extern "C" entry point catches these and turns them
// into a status code plus a thread-local message (mm_last_error()).
struct mm_error : public std::runtime_error {
    int code;
    mm_error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void mm_set_error(const std::string &msg);

// looks up, on the current device, the scratch array of the
// allocation-free entry point (mm_local.hip); called by mm_init
void mm_noalloc_setup();

#define MM_REQUIRE(cond, msg)                                                     \
    do {                                                                          \
        if (!(cond)) throw mm_error(MM_ERR_INVALID, std::string(__func__) + ": " + (msg)); \
    } while (0)

#define MM_CAPI_BEGIN try {
#define MM_CAPI_END                                                               \
    return MM_OK;                                                                 \
    }                                                                             \
    catch (const mm_error &e) {                                                   \
        mm_set_error(e.what());                                                   \
        return e.code;                                                            \
    }                                                                             \
    catch (const std::exception &e) {                                             \
        mm_set_error(e.what());                                                   \
        return MM_ERR_RUNTIME;                                                    \
    }                                                                             \
    catch (...) {                                                                 \
        mm_set_error("unknown C++ exception");                                    \
        return MM_ERR_RUNTIME;                                                    \
    }

// host-side tensor generator (demag_tensor.cpp)
int mm_round_tensor_dimension_impl(int dim, bool periodic, int padding);
void mm_demag_tensor_impl(int nx, int ny, int nz, double dx, double dy, double dz,
                          int ex, int ey, int ez, double *N, int nthreads,
                          int px, int py, int pz, int repeat);
void mm_demag_entry_open_ld(int i, int j, int k, double dx, double dy, double dz, double acc[6]);
void mm_demag_fold_open(int i, int j, int k, int ex, int ey, int ez, const double v[6], double *N);
void mm_demag_tensor_entry_impl(int i, int j, int k, double dx, double dy, double dz, double v6[6]);

// constants (src/magneto/mmm/constants.h:26-40)
static const long double MM_PI_L = 3.141592653589793238462643383279502884197169399L;
#define MM_MU0_VALUE ((double)(4.0L * MM_PI_L * (long double)1e-7))
static const double MM_MU_BOHR = 9.2740154e-24;
static const double MM_E_CHARGE = 1.602176487e-19;
static const double MM_GAMMA = 2.211e5;

#endif
