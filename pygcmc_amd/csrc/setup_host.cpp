// setup_host.cpp -- host-side system set-up in native code
// (gcmc/python/dataset.py:21-188).
//
// gcmc_grid_set_order: the reference lists the free cavity-grid cells in the
// iteration order of the Python set it builds with
//   {(x, y, z) for x in range(n0) for y in range(n1) for z in range(n2)}
// (dataset.py:66) and then discards the occupied cells from.  Discarding only
// marks slots, so the order is the slot order of the full set.  That order is
// fixed by CPython's tuple hash and its set table (open addressing: up to ten
// consecutive slots, then the perturbed step i = 5 i + 1 + (perturb >>= 5);
// growth once 5 fill >= 3 mask, to the smallest power of two above 4 used, or
// 2 used past 50,000 entries, re-inserting in old slot order).  This file
// computes the same slots without building Python objects.  The tuple hash
// is CPython 3.8+'s xxHash-derived combination of the item hashes, and a
// small non-negative int hashes to itself.
//
// gcmc_fixed_groups: consecutive grouping of the fixed atoms
// (dataset.py:175-188): an atom opens a new group when it lies more than the
// cutoff from the first atom of the current group, FP64, no periodic images.
// The squares are Python's float ** 2, i.e. the C library's pow(d, 2.0),
// which differs from d * d in the last bit for about one value in a thousand;
// the call goes through a volatile pointer so the compiler cannot turn it
// into a multiplication.
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/pygcmc_hip.h"

namespace gcmc_internal {
int set_error(int code, const std::string &msg);
}

namespace {

constexpr uint64_t XXPRIME_1 = 11400714785074694791ull;
constexpr uint64_t XXPRIME_2 = 14029467366897019727ull;
constexpr uint64_t XXPRIME_5 = 2870177450012600261ull;
constexpr int LINEAR_PROBES = 9;
constexpr int PERTURB_SHIFT = 5;
constexpr size_t SET_MINSIZE = 8;

// hash((x, y, z)) of non-negative ints below 2^61 - 1
uint64_t tuple3_hash(uint64_t x, uint64_t y, uint64_t z)
{
    uint64_t acc = XXPRIME_5;
    const uint64_t lanes[3] = {x, y, z};
    for (uint64_t lane : lanes) {
        acc += lane * XXPRIME_2;
        acc = (acc << 31) | (acc >> 33);
        acc *= XXPRIME_1;
    }
    acc += 3ull ^ (XXPRIME_5 ^ 3527539ull);
    if (acc == ~0ull) return 1546275796ull;
    return acc;
}

struct SetTable {
    std::vector<int64_t> key;   // linear cell index, -1 = empty
    std::vector<uint64_t> hash;
    size_t mask = SET_MINSIZE - 1;

    explicit SetTable(size_t size) : key(size, -1), hash(size, 0), mask(size - 1) {}

    // first empty slot of the probe sequence (keys are distinct, no dummies)
    void insert(int64_t k, uint64_t h)
    {
        size_t perturb = h;
        size_t i = h & mask;
        while (true) {
            if (key[i] < 0) {
                key[i] = k;
                hash[i] = h;
                return;
            }
            if (i + LINEAR_PROBES <= mask) {
                for (int j = 1; j <= LINEAR_PROBES; ++j) {
                    if (key[i + j] < 0) {
                        key[i + j] = k;
                        hash[i + j] = h;
                        return;
                    }
                }
            }
            perturb >>= PERTURB_SHIFT;
            i = (i * 5 + 1 + perturb) & mask;
        }
    }
};

}  // namespace

extern "C" int gcmc_grid_set_order(const int *grid_n, const unsigned char *occupied, long long max_cells, int *cells,
                                   long long *count)
{
    if (!grid_n || !count || (cells && max_cells < 0))
        return gcmc_internal::set_error(GCMC_E_ARG, "bad arguments");
    for (int a = 0; a < 3; ++a)
        if (grid_n[a] < 1) return gcmc_internal::set_error(GCMC_E_ARG, "grid_n must be positive");
    const int64_t n0 = grid_n[0], n1 = grid_n[1], n2 = grid_n[2];
    const int64_t total = n0 * n1 * n2;
    // the emulated table takes 16 B per slot, up to 4 slots per cell
    if (total > (1ll << 28)) return gcmc_internal::set_error(GCMC_E_ARG, "grid too large (more than 2^28 cells)");
    long long nfree = total;
    if (occupied) {
        nfree = 0;
        for (int64_t c = 0; c < total; ++c) nfree += occupied[c] ? 0 : 1;
    }
    *count = nfree;
    if (!cells) return 0;
    if (nfree > max_cells) return gcmc_internal::set_error(GCMC_E_ARG, "cell buffer too small for the free cells");

    SetTable t(SET_MINSIZE);
    size_t fill = 0, used = 0;
    int64_t c = 0;
    for (int64_t x = 0; x < n0; ++x)
        for (int64_t y = 0; y < n1; ++y)
            for (int64_t z = 0; z < n2; ++z, ++c) {
                t.insert(c, tuple3_hash((uint64_t)x, (uint64_t)y, (uint64_t)z));
                ++fill;
                ++used;
                if (fill * 5 < t.mask * 3) continue;
                const size_t minused = used > 50000 ? used * 2 : used * 4;
                size_t size = SET_MINSIZE;
                while (size <= minused) size <<= 1;
                SetTable g(size);
                for (size_t s = 0; s <= t.mask; ++s)
                    if (t.key[s] >= 0) g.insert(t.key[s], t.hash[s]);
                t = std::move(g);
                fill = used;
            }
    long long o = 0;
    for (size_t s = 0; s <= t.mask; ++s) {
        const int64_t k = t.key[s];
        if (k < 0 || (occupied && occupied[k])) continue;
        cells[3 * o + 0] = (int)(k / (n1 * n2));
        cells[3 * o + 1] = (int)((k / n2) % n1);
        cells[3 * o + 2] = (int)(k % n2);
        ++o;
    }
    return 0;
}

extern "C" int gcmc_fixed_groups(int n, const double *xyz, double cutoff, int *group)
{
    if (n < 0 || (n > 0 && (!xyz || !group))) return gcmc_internal::set_error(GCMC_E_ARG, "bad arguments");
    double (*volatile py_pow)(double, double) = static_cast<double (*)(double, double)>(&::pow);
    double x0 = INFINITY, y0 = INFINITY, z0 = INFINITY;
    int g = -1;
    for (int i = 0; i < n; ++i) {
        const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
        const double s = py_pow(x - x0, 2.0) + py_pow(y - y0, 2.0) + py_pow(z - z0, 2.0);
        if (std::sqrt(s) > cutoff) {
            ++g;
            x0 = x;
            y0 = y;
            z0 = z;
        }
        group[i] = g;
    }
    return 0;
}
