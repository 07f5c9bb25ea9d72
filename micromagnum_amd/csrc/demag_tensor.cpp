// demag_tensor.cpp -- demagnetisation tensor field for open boundaries.
//
// Host-side setup code (long double, x87 64-bit mantissa), run once per
// geometry.  Restates the reference generator
//   round_tensor_dimension     src/magneto/mmm/demag/tensor_round.cpp:27-106
//   calculateDemagTensor       src/magneto/mmm/demag/tensor.cpp:35-191
//   I_T / sumf / fseries /     src/magneto/mmm/demag/tensor_integrals.h:313-585
//   sumf_large_x
//   GenerateDemagTensor (N:=-N) src/magneto/mmm/demag/demag_tensor.cpp:116-186
// with the generalized Newell functions taken from newell_table.inc, which
// gen_newell_table.py derives symbolically from Newell's f and g.
//
// Per entry (i,j,k) in [0,n)^3 and component (xx,xy,xz,yy,yz,zz):
//   * axes are sorted so that the cell edges satisfy lx <= ly <= lz;
//   * entries odd in an axis vanish on that axis' zero plane;
//   * near field: 27-point (3 s^2 - 2)-weighted second difference of the
//     summands of Newell's f (diagonal) or g (off-diagonal);
//   * points far along one axis inside a narrow cone: asymptotic large-x sum;
//   * axes whose edge is small against the distance (radius 30 edges):
//     Taylor series of the second difference in that axis (centre or, on the
//     zero plane, border expansion);
//   * result mirrored into the 8 octants of the padded (ex, ey, ez) grid with
//     the sign pattern of each component.
// The returned field is the one convolved with M (sign already flipped).

#include "mm_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <functional>
#include <stdexcept>
#include <thread>
#include <vector>

namespace {

typedef long double ld;

// pi as the reference declares it (src/magneto/mmm/constants.h:26): a double
// literal stored in a long double, i.e. pi rounded to 53 bits
static const ld PI_L = (ld)3.141592653589793238462643383279502884197169399;
static const double EXPANSION_RADIUS = 30.0;
static const double CONE_RATIO = 6.0;

#include "newell_table.inc"
#include "pbc_table.inc"

// ---------------------------------------------------------------------------
// padding policy (tensor_round.cpp:27-106)
// ---------------------------------------------------------------------------
int next_multiple(int x, int m) { return (x % m) ? (x / m + 1) * m : x; }

int next_pow2(int x)
{
    int p = 1;
    while (p < x) p <<= 1;
    return p;
}

int largest_prime_factor(int n)
{
    if (n <= 1) return 1;
    int best = 1;
    for (int f = 2; (long)f * f <= n; ++f)
        while (n % f == 0) { best = f; n /= f; }
    return n > 1 ? std::max(best, n) : best;
}

// ---------------------------------------------------------------------------
// f(a,b,c) for arbitrary order: sort so that a <= b <= c, permuting the axes
// (tensor_integrals.h:113-118 behaviour), then evaluate the table entry.
// ---------------------------------------------------------------------------
ld newell_f(int a, int b, int c, ld lx, ld ly, ld lz, long i, long j, long k)
{
    if (a > b) return newell_f(b, a, c, ly, lx, lz, j, i, k);
    if (b > c) return newell_f(a, c, b, lx, lz, ly, i, k, j);
    ld out = 0.0L;
    if (!newell_f_sorted(a, b, c, (ld)i * lx, (ld)j * ly, (ld)k * lz, &out))
        throw std::runtime_error("demag tensor: generalized Newell function not tabulated");
    return out;
}

// Near-field summands.  Each summand of Newell's f (n = 0..7) and g
// (n = 0..8) is evaluated on its own; a summand whose polynomial prefactor is
// zero is returned as zero, which is its limit and also skips atan(./0) and
// atanh(+-1).
ld f002_term(int n, ld x, ld y, ld z)
{
    const ld R = sqrtl(x * x + y * y + z * z);
    switch (n) {
    case 0: return -R * x * x / 6.0L;
    case 1: return -R * y * y / 6.0L;
    case 2: return R * z * z / 3.0L;
    case 3: return (x == 0 || y == 0 || z == 0) ? 0.0L : -x * y * z * atanl((x * y) / (R * z));
    case 4: return (x == 0 || y == 0) ? 0.0L : x * y * y * atanhl(x / R) / 2.0L;
    case 5: return (x == 0 || z == 0) ? 0.0L : -x * z * z * atanhl(x / R) / 2.0L;
    case 6: return (x == 0 || y == 0) ? 0.0L : x * x * y * atanhl(y / R) / 2.0L;
    case 7: return (y == 0 || z == 0) ? 0.0L : -y * z * z * atanhl(y / R) / 2.0L;
    default: return 0.0L;
    }
}

ld f011_term(int n, ld x, ld y, ld z)
{
    const ld R = sqrtl(x * x + y * y + z * z);
    switch (n) {
    case 0: return -R * y * z / 3.0L;
    case 1: return (x == 0 || y == 0 || z == 0) ? 0.0L : -x * z * z * atanl((x * y) / (R * z)) / 2.0L;
    case 2: return (x == 0 || y == 0 || z == 0) ? 0.0L : -x * y * y * atanl((x * z) / (R * y)) / 2.0L;
    case 3: return (x == 0 || y == 0 || z == 0) ? 0.0L : -x * x * x * atanl((y * z) / (R * x)) / 6.0L;
    case 4: return (x == 0 || y == 0 || z == 0) ? 0.0L : x * y * z * atanhl(x / R);
    case 5: return (x == 0 || z == 0) ? 0.0L : x * x * z * atanhl(y / R) / 2.0L;
    case 6: return (z == 0) ? 0.0L : -z * z * z * atanhl(y / R) / 6.0L;
    case 7: return (x == 0 || y == 0) ? 0.0L : x * x * y * atanhl(z / R) / 2.0L;
    case 8: return (y == 0) ? 0.0L : -y * y * y * atanhl(z / R) / 6.0L;
    default: return 0.0L;
    }
}

// summand n of the base function for component order (a,b,c) (any order of
// a permutation of (0,0,2) or (0,1,1)), axes permuted like newell_f.
ld base_term(int a, int b, int c, ld lx, ld ly, ld lz, long i, long j, long k, int n)
{
    if (a > b) return base_term(b, a, c, ly, lx, lz, j, i, k, n);
    if (b > c) return base_term(a, c, b, lx, lz, ly, i, k, j, n);
    const ld x = (ld)i * lx, y = (ld)j * ly, z = (ld)k * lz;
    if (a == 0 && b == 0 && c == 2) return f002_term(n, x, y, z);
    if (a == 0 && b == 1 && c == 1) return f011_term(n, x, y, z);
    throw std::runtime_error("demag tensor: unsupported base function");
}

inline ld w3(int s) { return (ld)(3 * s * s - 2); }

double factorial(int n)
{
    double r = 1.0;
    for (int q = 2; q <= n; ++q) r *= q;
    return r;
}

ld ipow_l(ld b, int n)
{
    ld r = 1.0L;
    for (int q = 0; q < n; ++q) r *= b;
    return r;
}

enum { AT_CENTER = 0, AT_BORDER = 1 };

// Taylor expansion of the second difference along x to order O; P, Q are
// either expansion orders along y, z or 0 (exact second difference there).
ld fseries(int o, int p, int q, int O, int P, int Q, ld lx, ld ly, ld lz,
           long i, long j, long k, int where)
{
    ld xp = 1.0L;
    if (where == AT_CENTER) {
        xp *= 2.0L * ipow_l(lx, O) / (ld)factorial(O);
    } else {
        xp *= (ld)((1 << O) - 2) * ipow_l(-lx, O) / (ld)factorial(O);
        i += 1; // expansion point shifted one cell out
    }
    if (Q != 0) {
        return xp * 2.0L * ipow_l(ly, P) / (ld)factorial(P) * 2.0L * ipow_l(lz, Q) / (ld)factorial(Q)
               * newell_f(o + O, p + P, q + Q, lx, ly, lz, i, j, k);
    }
    if (P != 0) {
        ld s1 = 0.0L;
        for (int s = -1; s <= 1; ++s) s1 += w3(s) * newell_f(o + O, p + P, q, lx, ly, lz, i, j, k + s);
        return xp * 2.0L * ipow_l(ly, P) / (ld)factorial(P) * s1;
    }
    ld s1 = 0.0L;
    for (int r = -1; r <= 1; ++r) {
        ld s2 = 0.0L;
        for (int s = -1; s <= 1; ++s) s2 += w3(s) * newell_f(o + O, p, q, lx, ly, lz, i, j + r, k + s);
        s1 += w3(r) * s2;
    }
    return xp * s1;
}

// asymptotic sum for points far out along x, components (0,0,2)/(0,2,0)/(0,1,1)
ld large_x(int o, int p, int q, ld lx, ld ly, ld lz, long i, long j, long k)
{
    if (p > q) return large_x(o, q, p, lx, lz, ly, i, k, j);
    const ld x = (ld)i * lx, y = (ld)j * ly, z = (ld)k * lz;
    double s;
    if (o == 0 && p == 0 && q == 2) s = (double)nlarge_002(lx, ly, lz, x, y, z);
    else if (o == 0 && p == 1 && q == 1) s = (double)nlarge_011(lx, ly, lz, x, y, z);
    else throw std::runtime_error("demag tensor: unsupported large-x component");
    return (ld)s;
}

long min_sq(long i) { return std::min(std::min((i + 1) * (i + 1), (i - 1) * (i - 1)), i * i); }

// second difference of the base function, cell edges (1, ly, lz), 1 <= ly <= lz
ld second_difference(int o, int p, int q, ld ly, ld lz, long i, long j, long k)
{
    const ld lx = 1.0L;
    const double er = EXPANSION_RADIUS;
    const ld Rsq = (ld)min_sq(i) * lx * lx + (ld)min_sq(j) * ly * ly + (ld)min_sq(k) * lz * lz;
    const ld lim = Rsq / (er * er);
    int expand = 0;
    if (lx * lx < lim) ++expand;
    if (ly * ly < lim) ++expand;
    if (lz * lz < lim) ++expand;

    const ld cone = Rsq / (CONE_RATIO * CONE_RATIO);
    const long ai = std::labs(i) + 1, aj = std::labs(j) + 1, ak = std::labs(k) + 1;
    const bool nx0 = (ld)(ai * ai) * lx * lx < cone;
    const bool ny0 = (ld)(aj * aj) * ly * ly < cone;
    const bool nz0 = (ld)(ak * ak) * lz * lz < cone;
    if (ny0 && nz0 && ((o == 0 && p == 2 && q == 0) || (o == 0 && p == 0 && q == 2) || (o == 0 && p == 1 && q == 1)))
        return large_x(o, p, q, lx, ly, lz, i, j, k);
    if (nx0 && nz0 && ((o == 2 && p == 0 && q == 0) || (o == 0 && p == 0 && q == 2) || (o == 1 && p == 0 && q == 1)))
        return large_x(p, o, q, ly, lx, lz, j, i, k);
    if (nx0 && ny0 && ((o == 2 && p == 0 && q == 0) || (o == 0 && p == 2 && q == 0) || (o == 1 && p == 1 && q == 0)))
        return large_x(q, o, p, lz, lx, ly, k, i, j);

    ld sum = 0.0L;
    if (expand == 3) {
        sum += fseries(o, p, q, 2, 2, 2, lx, ly, lz, i, j, k, AT_CENTER);
        sum += fseries(o, p, q, 4, 2, 2, lx, ly, lz, i, j, k, AT_CENTER);
        sum += fseries(o, p, q, 2, 4, 2, lx, ly, lz, i, j, k, AT_CENTER);
        sum += fseries(o, p, q, 2, 2, 4, lx, ly, lz, i, j, k, AT_CENTER);
    } else if (expand == 2) {
        if (i == 0 && j == 0) {
            for (int O = 2; O <= 6; ++O) sum += fseries(o, p, q, O, 2, 0, lx, ly, lz, i, j, k, AT_BORDER);
            sum += fseries(o, p, q, 2, 4, 0, lx, ly, lz, i, j, k, AT_BORDER);
        } else {
            sum += fseries(o, p, q, 2, 2, 0, lx, ly, lz, i, j, k, AT_CENTER);
            sum += fseries(o, p, q, 4, 2, 0, lx, ly, lz, i, j, k, AT_CENTER);
            sum += fseries(o, p, q, 2, 4, 0, lx, ly, lz, i, j, k, AT_CENTER);
        }
    } else if (expand == 1) {
        if (i == 0) {
            for (int O = 2; O <= 7; ++O) sum += fseries(o, p, q, O, 0, 0, lx, ly, lz, i, j, k, AT_BORDER);
        } else {
            for (int O = 2; O <= 6; O += 2) sum += fseries(o, p, q, O, 0, 0, lx, ly, lz, i, j, k, AT_CENTER);
        }
    } else {
        for (int n = 0; n < 13; ++n) {
            ld s1 = 0.0L;
            for (int c = -1; c <= 1; ++c) {
                ld s2 = 0.0L;
                for (int b = -1; b <= 1; ++b) {
                    ld s3 = 0.0L;
                    for (int a = -1; a <= 1; ++a)
                        s3 += w3(a) * base_term(o, p, q, lx, ly, lz, i + a, j + b, k + c, n);
                    s2 += w3(b) * s3;
                }
                s1 += w3(c) * s2;
            }
            sum += s1;
        }
    }
    return sum;
}

// averaged-cell interaction integral of component (o,p,q) at offset (i,j,k)
ld cell_integral(int o, int p, int q, ld lx, ld ly, ld lz, long i, long j, long k)
{
    if (lx > ly) return cell_integral(p, o, q, ly, lx, lz, j, i, k);
    if (ly > lz) return cell_integral(o, q, p, lx, lz, ly, i, k, j);
    if ((o & 1) && i == 0) return 0.0L;
    if ((p & 1) && j == 0) return 0.0L;
    if ((q & 1) && k == 0) return 0.0L;
    const ld pref = ipow_l(lx, 5 - o - p - q) / (4.0L * PI_L * lx * ly * lz);
    return pref * second_difference(o, p, q, ly / lx, lz / lx, i, j, k);
}

} // namespace

int mm_round_tensor_dimension_impl(int dim, bool periodic, int padding)
{
    if (periodic) return dim;
    const int t = 2 * dim - 1;
    if (t == 1) return 1;
    switch (padding) {
    case MM_PADDING_DISABLE: return t;
    case MM_PADDING_ROUND_2: return next_multiple(t, 2);
    case MM_PADDING_ROUND_4: return next_multiple(t, 4);
    case MM_PADDING_ROUND_8: return next_multiple(t, 8);
    case MM_PADDING_ROUND_POT: return next_pow2(t);
    case MM_PADDING_SMALL_PRIME_FACTORS: {
        int v = t + 1;
        while (largest_prime_factor(v) > 5) ++v;
        return v;
    }
    default: throw std::runtime_error("Invalid padding strategy");
    }
}

namespace {

// Far-image sums of the point-dipole tensor (pbc_table.inc) for one periodic
// axis p (1-D) or two periodic axes p, s (2-D), the reference's PBC_Demag_1D /
// PBC_Demag_2D (tensor_integrals.h:464-519) called as in tensor.cpp:66-160:
// the periodic axes come first, `ord` holds the derivative order of the
// component along each of the three axes of that order.
static ld pbc1(const int ord[3], const ld l[3], const long idx[3], int np)
{
    const ld L = l[0] * (ld)np;
    const ld x = (ld)idx[0] * l[0] - 0.5L * L, y = (ld)idx[1] * l[1], z = (ld)idx[2] * l[2];
    return pbc1_sorted(ord[0], ord[1], ord[2], x, y, z, L);
}

static ld pbc2(const int ord[3], const ld l[3], const long idx[3], int n0, int n1)
{
    const ld L0 = l[0] * (ld)n0, L1 = l[1] * (ld)n1;
    const ld x = (ld)idx[0] * l[0] - 0.5L * L0, y = (ld)idx[1] * l[1] - 0.5L * L1, z = (ld)idx[2] * l[2];
    return pbc2_sorted(ord[0], ord[1], ord[2], x, y, z, L0, L1);
}

} // namespace

void mm_demag_tensor_impl(int nx, int ny, int nz, double dx, double dy, double dz,
                          int ex, int ey, int ez, double *N, int nthreads,
                          int px, int py, int pz, int repeat)
{
    const int n3[3] = {nx, ny, nz}, e3[3] = {ex, ey, ez}, per[3] = {px != 0, py != 0, pz != 0};
    for (int a = 0; a < 3; ++a) {
        if (per[a] ? e3[a] != n3[a] : e3[a] < 2 * n3[a] - 1)
            throw std::runtime_error(per[a] ? "demag tensor: a periodic axis needs the unpadded size"
                                            : "demag tensor: padded size too small for open boundaries");
    }
    if ((px || py || pz) && repeat < 1) throw std::runtime_error("demag tensor: periodic repeat must be >= 1");
    // images summed cell by cell per axis (demag_tensor.cpp:155-157)
    const int rep3[3] = {px ? repeat : 1, py ? repeat : 1, pz ? repeat : 1};
    const bool no_inf = std::getenv("MAGNUM_DEMAG_NO_INFINITY_CORRECTION") != nullptr;
    const size_t tot = (size_t)6 * ex * ey * ez;
    std::fill(N, N + tot, 0.0);
    const ld l3[3] = {(ld)dx, (ld)dy, (ld)dz};
    const ld lx = l3[0], ly = l3[1], lz = l3[2];
    // component e = (xx, xy, xz, yy, yz, zz) -> its two axes
    static const int CA[6] = {0, 0, 0, 1, 1, 2}, CB[6] = {0, 1, 2, 1, 2, 2};
    const double scale = (double)(-lx * ly * lz / (4.0L * PI_L));

    // the six entries for offset (i,j,k) before mirroring, as the reference
    // accumulates them: double sums, each term rounded once
    auto entry = [&](int i, int j, int k, double *v) {
        double acc[6] = {0, 0, 0, 0, 0, 0};
        for (int rx = 0; rx < rep3[0]; ++rx)
            for (int ry = 0; ry < rep3[1]; ++ry)
                for (int rz = 0; rz < rep3[2]; ++rz) {
                    const long I = i + (long)rx * nx, J = j + (long)ry * ny, K = k + (long)rz * nz;
                    int mcs = 1;
                    if (I == 0) mcs *= 2;
                    if (J == 0) mcs *= 2;
                    if (K == 0) mcs *= 2;
                    for (int e = 0; e < 6; ++e) {
                        int o[3] = {0, 0, 0};
                        ++o[CA[e]];
                        ++o[CB[e]];
                        acc[e] = (double)((ld)acc[e] - cell_integral(o[0], o[1], o[2], lx, ly, lz, I, J, K) / (ld)mcs);
                    }
                }
        if (!no_inf) {
            const int idx0[3] = {i, j, k};
            // one periodic axis p: tails beyond the repeat zone for every image
            // of the other axes inside theirs (tensor.cpp:66-118)
            for (int p = 0; p < 3; ++p) {
                if (rep3[p] <= 1) continue;
                const int s = p == 0 ? 1 : 0, t = p == 2 ? 1 : 2;
                const ld l[3] = {l3[p], l3[s], l3[t]};
                for (int rs = 0; rs < rep3[s]; ++rs)
                    for (int rt = 0; rt < rep3[t]; ++rt) {
                        const long idx[3] = {idx0[p] + (long)rep3[p] * n3[p], idx0[s] + (long)rs * n3[s],
                                             idx0[t] + (long)rt * n3[t]};
                        int mcs = 1;
                        if (idx[1] == 0) mcs *= 2;
                        if (idx[2] == 0) mcs *= 2;
                        for (int e = 0; e < 6; ++e) {
                            int oa[3] = {0, 0, 0};
                            ++oa[CA[e]];
                            ++oa[CB[e]];
                            const int ord[3] = {oa[p], oa[s], oa[t]};
                            acc[e] += scale * (double)pbc1(ord, l, idx, n3[p]) / mcs;
                        }
                    }
            }
            // two periodic axes p < s: quadrant tails (tensor.cpp:120-160)
            for (int p = 0; p < 3; ++p)
                for (int s = p + 1; s < 3; ++s) {
                    if (rep3[p] <= 1 || rep3[s] <= 1) continue;
                    const int t = 3 - p - s;
                    const ld l[3] = {l3[p], l3[s], l3[t]};
                    for (int rt = 0; rt < rep3[t]; ++rt) {
                        const long idx[3] = {idx0[p] + (long)rep3[p] * n3[p], idx0[s] + (long)rep3[s] * n3[s],
                                             idx0[t] + (long)rt * n3[t]};
                        const int mcs = idx[2] == 0 ? 2 : 1;
                        for (int e = 0; e < 6; ++e) {
                            int oa[3] = {0, 0, 0};
                            ++oa[CA[e]];
                            ++oa[CB[e]];
                            const int ord[3] = {oa[p], oa[s], oa[t]};
                            acc[e] += scale * (double)pbc2(ord, l, idx, n3[p], n3[s]) / mcs;
                        }
                    }
                }
        }
        for (int e = 0; e < 6; ++e) v[e] = acc[e];
    };

    // mirror into the 8 octants with the sign pattern of each component
    // (tensor.cpp:167-180); along a periodic axis -i wraps to n - i
    auto fold = [&](int i, int j, int k, const double *v) {
        for (int o = -1; o <= 1; o += 2)
            for (int p = -1; p <= 1; p += 2)
                for (int q = -1; q <= 1; q += 2) {
                    const int I = (o * i + ex) % ex, J = (p * j + ey) % ey, K = (q * k + ez) % ez;
                    double *e = N + 6 * ((size_t)I + (size_t)ex * ((size_t)J + (size_t)ey * K));
                    e[0] += v[0];
                    e[1] += o * p * v[1];
                    e[2] += o * q * v[2];
                    e[3] += v[3];
                    e[4] += p * q * v[4];
                    e[5] += v[5];
                }
    };

    if (nthreads <= 0) {
        const char *env = std::getenv("MM_TENSOR_THREADS");
        nthreads = env ? std::atoi(env) : (int)std::thread::hardware_concurrency();
        if (nthreads <= 0) nthreads = 1;
    }
    nthreads = std::max(1, std::min(nthreads, nx));
    auto parallel_over_i = [&](const std::function<void(int)> &f) {
        if (nthreads <= 1) {
            for (int i = 0; i < nx; ++i) f(i);
            return;
        }
        std::vector<std::thread> pool;
        for (int t = 0; t < nthreads; ++t)
            pool.emplace_back([&, t]() {
                for (int i = t; i < nx; i += nthreads) f(i);
            });
        for (auto &th : pool) th.join();
    };

    if (!px) {
        // distinct i map to distinct x-planes of the padded grid (ex >= 2nx-1),
        // so threads never write the same entry and every entry sees its
        // terms in the reference's order
        parallel_over_i([&](int i) {
            double v[6];
            for (int j = 0; j < ny; ++j)
                for (int k = 0; k < nz; ++k) {
                    entry(i, j, k, v);
                    fold(i, j, k, v);
                }
        });
    } else {
        // i and nx - i share a plane: compute the entries in parallel, then
        // mirror them in the reference's loop order
        const size_t n = (size_t)nx * ny * nz;
        std::vector<double> vals(6 * n);
        parallel_over_i([&](int i) {
            for (int j = 0; j < ny; ++j)
                for (int k = 0; k < nz; ++k) entry(i, j, k, &vals[6 * ((size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k))]);
        });
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j)
                for (int k = 0; k < nz; ++k) fold(i, j, k, &vals[6 * ((size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k))]);
    }
    for (size_t q = 0; q < tot; ++q) N[q] = -N[q];
}

// One open-boundary entry before mirroring, as `entry` above computes it
// without periodic images: acc[e] = -I_e(i,j,k) / mcs rounded to double.
void mm_demag_entry_open_ld(int i, int j, int k, double dx, double dy, double dz, double acc[6])
{
    static const int CA[6] = {0, 0, 0, 1, 1, 2}, CB[6] = {0, 1, 2, 1, 2, 2};
    int mcs = 1;
    if (i == 0) mcs *= 2;
    if (j == 0) mcs *= 2;
    if (k == 0) mcs *= 2;
    for (int e = 0; e < 6; ++e) {
        int o[3] = {0, 0, 0};
        ++o[CA[e]];
        ++o[CB[e]];
        acc[e] = (double)(0.0L - cell_integral(o[0], o[1], o[2], (ld)dx, (ld)dy, (ld)dz, i, j, k) / (ld)mcs);
    }
}

// Mirror one entry into the 8 octants of the padded grid, accumulating in
// the loop order of `fold` above (tensor.cpp:167-180), open boundaries.
void mm_demag_fold_open(int i, int j, int k, int ex, int ey, int ez, const double v[6], double *N)
{
    for (int o = -1; o <= 1; o += 2)
        for (int p = -1; p <= 1; p += 2)
            for (int q = -1; q <= 1; q += 2) {
                const int I = (o * i + ex) % ex, J = (p * j + ey) % ey, K = (q * k + ez) % ez;
                double *e = N + 6 * ((size_t)I + (size_t)ex * ((size_t)J + (size_t)ey * K));
                e[0] += v[0];
                e[1] += o * p * v[1];
                e[2] += o * q * v[2];
                e[3] += v[3];
                e[4] += p * q * v[4];
                e[5] += v[5];
            }
}

// The padded open-boundary tensor's value at (i, j, k) (0 <= i < nx, ...):
// the mirror images of entry (i, j, k) that land there, accumulated in the
// loop order of `fold` and sign-flipped, i.e. what mm_demag_tensor_impl
// stores at that location (for any padding >= 2n - 1).
void mm_demag_tensor_entry_impl(int i, int j, int k, double dx, double dy, double dz, double v6[6])
{
    double acc[6];
    mm_demag_entry_open_ld(i, j, k, dx, dy, dz, acc);
    double e[6] = {0, 0, 0, 0, 0, 0};
    for (int o = -1; o <= 1; o += 2)
        for (int p = -1; p <= 1; p += 2)
            for (int q = -1; q <= 1; q += 2) {
                if ((o < 0 && i != 0) || (p < 0 && j != 0) || (q < 0 && k != 0)) continue;
                e[0] += acc[0];
                e[1] += o * p * acc[1];
                e[2] += o * q * acc[2];
                e[3] += acc[3];
                e[4] += p * q * acc[4];
                e[5] += acc[5];
            }
    for (int c = 0; c < 6; ++c) v6[c] = -e[c];
}
