// ihg_group.hip -- the batched form across grid sizes: every pass of mk_rhs
// and of the RK4 stages runs as ONE launch over all realizations of several
// ensemble contexts of different grid sizes (a group).  No counterpart in
// the reference, which runs one realization per process.
//
// A grouped launch is a one-dimensional grid over a table of segments, one
// per context.  Segment s owns the work-groups [begin_s, begin_s + gx gy gz)
// -- exactly the grid its context's own launch of the pass would have, laid
// out x fastest -- and begin_s is a multiple of 8, so that work-group w of a
// segment lands on XCD w mod 8 as in the standalone launch and the
// XCD-aware tile order keeps its meaning.  Segments are ordered by grid
// size, largest first, so that the longest work-groups start first.  Every
// work-group finds its segment with at most GMAX - 1 uniform compares and
// runs the pass's device body (fft_kernels.h, rhs_kernels.h, rk4_kernels.h)
// for that segment's line length: the bodies are the ones the per-context
// kernels run, so a grouped step gives the per-context step's values bit for
// bit.  The table travels in the kernel arguments, so a captured graph of a
// grouped step holds everything it needs.
#include <algorithm>
#include <cstring>
#include <vector>

#include "dispatch.h"
#include "fft_kernels.h"
#include "ihg_internal.h"
#include "mem.h"
#include "reduce.h"
#include "rhs_kernels.h"
#include "rk4_kernels.h"

using namespace ihg;

namespace ihg {

constexpr int GMAX = 4; // contexts per group

// line lengths (cubic grids) a grouped launch serves
constexpr int GLENS[] = {16, 32, 64, 128};
constexpr int NGLENS = sizeof(GLENS) / sizeof(GLENS[0]);

static bool group_len_ok(int L)
{
    for (int i = 0; i < NGLENS; ++i)
        if (GLENS[i] == L) return true;
    return false;
}

struct SegGeom {
    unsigned begin;       // first work-group (a multiple of 8)
    unsigned gx, gy, gz;  // the segment's own grid
    int L;                // its line length
};

template <class A>
struct GroupTab {
    int n;
    SegGeom g[GMAX];
    A a[GMAX];
};

// LDS (complex values) a family needs for one line length, and the most
// over the served lengths; a grouped launch asks for the most over the
// lengths it holds (dynamic LDS), so a group without 16- or 128-point
// grids keeps the smaller footprint of its own lengths
template <class Fam>
static int fam_lds(int L)
{
    switch (L) {
    case 16: return Fam::template lds<16>();
    case 32: return Fam::template lds<32>();
    case 64: return Fam::template lds<64>();
    default: return Fam::template lds<128>();
    }
}

template <class Fam>
constexpr int fam_lds_max()
{
    return std::max({Fam::template lds<16>(), Fam::template lds<32>(),
                     Fam::template lds<64>(), Fam::template lds<128>()});
}

// LDS of the strided and fan-out bodies: none for one-stage plans, whose
// only stage loads from and stores to registers
template <int L>
constexpr int strided_lds()
{
    return Plan<L>::R1 > 1 ? StridedShape<L>::LDS : 0;
}

// Sets of line lengths a grouped kernel is compiled for (bit i: GLENS[i]).
// A kernel holds the bodies of its set only, so its register count is the
// largest of those bodies: the set of the mixed ensemble (32, 64, 128)
// leaves out the 16-point bodies, whose 16 values per thread would cost the
// other lengths occupancy.  A group runs the smallest compiled set that
// holds its lengths.
constexpr int LEN_BIT[] = {1, 2, 4, 8};
constexpr int SET_ALL = 15, SET_32_128 = 14;

static int len_bit(int L)
{
    for (int i = 0; i < NGLENS; ++i)
        if (GLENS[i] == L) return LEN_BIT[i];
    return 0;
}

template <class Fam, int SET>
__device__ __forceinline__ void run_len(int L, const Blk &bk, double2 *lds,
                                        const typename Fam::Args &a)
{
    if constexpr ((SET & 1) != 0)
        if (L == 16) { Fam::template run<16>(bk, lds, a); return; }
    if constexpr ((SET & 2) != 0)
        if (L == 32) { Fam::template run<32>(bk, lds, a); return; }
    if constexpr ((SET & 4) != 0)
        if (L == 64) { Fam::template run<64>(bk, lds, a); return; }
    if constexpr ((SET & 8) != 0)
        if (L == 128) { Fam::template run<128>(bk, lds, a); return; }
}

template <class Fam, int SET>
__global__ __launch_bounds__(Fam::WG) void k_group(
    GroupTab<typename Fam::Args> t)
{
    extern __shared__ double2 lds[];
    const unsigned bid = blockIdx.x;
    int s = 0;
#pragma unroll
    for (int i = 1; i < GMAX; ++i)
        if (i < t.n && bid >= t.g[i].begin) s = i;
    const SegGeom g = t.g[s];
    const unsigned w = bid - g.begin, per_z = g.gx * g.gy;
    if (w >= per_z * g.gz) return; // alignment padding: whole work-groups
    const Blk bk{w % g.gx, (w / g.gx) % g.gy, w / per_z, g.gx};
    if constexpr (Fam::BY_LEN)
        run_len<Fam, SET>(g.L, bk, lds, t.a[s]);
    else
        Fam::template run<0>(bk, lds, t.a[s]);
}

// ---------------------------------------------------------------------------
// families: the argument record of one segment and the body it runs
// ---------------------------------------------------------------------------

struct FamZR2C {
    static constexpr int WG = FFT_WG;
    static constexpr bool BY_LEN = true;
    struct Args {
        const double *in;
        double2 *out;
        long nrows, rbatch, cbatch;
        const double2 *tw;
        long rin, rout;
    };
    template <int L> static constexpr int lds() { return Tile<L>::B * Tile<L>::PITCH; }
    template <int L>
    __device__ static void run(const Blk &bk, double2 *lds, const Args &a)
    {
        z_r2c_body<L>(bk, lds, a.in, a.out, a.nrows, a.rbatch, a.cbatch, a.tw,
                      a.rin, a.rout);
    }
};

struct FamZC2R {
    static constexpr int WG = FFT_WG;
    static constexpr bool BY_LEN = true;
    struct Args {
        const double2 *in;
        double *out;
        long nrows, cbatch, rbatch;
        const double2 *tw;
        long rin, rout;
    };
    template <int L> static constexpr int lds() { return Tile<L>::B * Tile<L>::PITCH; }
    template <int L>
    __device__ static void run(const Blk &bk, double2 *lds, const Args &a)
    {
        z_c2r_body<L>(bk, lds, a.in, a.out, a.nrows, a.cbatch, a.rbatch, a.tw,
                      a.rin, a.rout);
    }
};

// y and x passes with the default line plan of every served length
// (Plan<L>, 256 lanes, XCD-aware tile order: candidate 0 of StridedCands)
template <int DIR>
struct FamStrided {
    static constexpr int WG = FFT_WG;
    static constexpr bool BY_LEN = true;
    struct Args {
        const double2 *in;
        double2 *out;
        LineSet ls;
        long batch;
        const double2 *tw;
        long rr;
    };
    template <int L> static constexpr int lds() { return strided_lds<L>(); }
    template <int L>
    __device__ static void run(const Blk &bk, double2 *lds, const Args &a)
    {
        c2c_strided_body<L, DIR>(bk, lds, a.in, a.out, a.ls, a.batch, a.tw,
                                 a.rr);
    }
};

template <int NQ, bool FWD, class Fac>
struct FamFanout {
    static constexpr int WG = FFT_WG;
    static constexpr bool BY_LEN = true;
    struct Args {
        const double2 *in;
        double2 *out;
        long qstride;
        LineSet ls;
        long batch;
        const double2 *tw;
        Fac fac;
        long rr;
    };
    template <int L> static constexpr int lds() { return strided_lds<L>(); }
    template <int L>
    __device__ static void run(const Blk &bk, double2 *lds, const Args &a)
    {
        x_fanout_body<L, NQ, FWD, Fac, typename FanPlan<L>::P, FanPlan<L>::WG>(
            bk, lds, a.in, a.out, a.qstride, a.ls, a.batch, a.tw, a.fac, a.rr);
    }
};

struct FamZC2RRho {
    static constexpr int WG = FFT_WG;
    static constexpr bool BY_LEN = true;
    struct Args {
        const double2 *W;
        long M;
        const double *f;
        double *result, *grad3, *partials;
        long nrows, N;
        double m2;
        const double2 *tw;
        Ens e;
    };
    template <int L> static constexpr int lds() { return z_rho_lds<L>(); }
    template <int L>
    __device__ static void run(const Blk &bk, double2 *lds, const Args &a)
    {
        z_c2r_rho_body<L, double>(bk, lds, a.W, a.M, a.f, a.result, a.grad3,
                                  nullptr, nullptr, a.partials, a.nrows, a.N,
                                  a.m2, a.tw, a.e);
    }
};

template <bool EP>
struct FamMeans {
    static constexpr int WG = RED_WG;
    static constexpr bool BY_LEN = false;
    struct Args {
        const double *partials;
        int nparts;
        long N;
        const double *f;
        double *result;
        long Ntot;
        double *dscal;
        Ens e;
        Rk4Ep ep;
    };
    template <int L> static constexpr int lds() { return 4; } // 8 doubles
    template <int L>
    __device__ static void run(const Blk &bk, double2 *lds, const Args &a)
    {
        // one work-group per realization: x is the realization
        means_body<EP>(bk.x, reinterpret_cast<double *>(lds), a.partials,
                       a.nparts, a.N, a.f, a.result, a.Ntot, a.dscal, a.e,
                       a.ep);
    }
};

template <bool EP>
struct FamUpdate {
    static constexpr int WG = RED_WG;
    static constexpr bool BY_LEN = false;
    struct Args {
        const double *f;
        double *result;
        long N, Ntot;
        double m2;
        const double *dscal;
        Ens e;
        int nt;
        Rk4Ep ep;
    };
    template <int L> static constexpr int lds() { return 0; }
    template <int L>
    __device__ static void run(const Blk &bk, double2 *, const Args &a)
    {
        if (a.nt)
            update_phi_psi_body<true, EP>(bk.x, bk.gx, bk.z, a.f, a.result, a.N,
                                          a.Ntot, a.m2, a.dscal, a.e, a.ep);
        else
            update_phi_psi_body<false, EP>(bk.x, bk.gx, bk.z, a.f, a.result,
                                           a.N, a.Ntot, a.m2, a.dscal, a.e,
                                           a.ep);
    }
};

struct FamSources {
    static constexpr int WG = FFT_WG;
    static constexpr bool BY_LEN = true;
    struct Args {
        const double *g3;
        double2 *W;
        long M, nrows, N;
        const double2 *tw;
        long rg, rw;
    };
    template <int L> static constexpr int lds() { return Tile<L>::B * Tile<L>::PITCH; }
    template <int L>
    __device__ static void run(const Blk &bk, double2 *lds, const Args &a)
    {
        z_r2c_sources_body<L, double2>(bk, lds, a.g3, a.W, a.M, a.nrows, a.N,
                                       a.tw, a.rg, a.rw);
    }
};

// plans of the fused TT pass at 256 lanes (the per-context pass runs 128-
// point lines with 512 lanes; the plan, and so every value, is the same)
template <int L> struct GroupTTPlan { using P = typename TTPlan<L>::P; };

template <bool EP>
struct FamTT {
    static constexpr int WG = FFT_WG;
    static constexpr bool BY_LEN = true;
    struct Args {
        const double2 *W;
        long M;
        const double *f;
        double *result;
        long N, Next, Ntot;
        KGeom kg;
        const double *dscal;
        LineSet ls;
        const double2 *tw;
        Ens e;
        long rw;
        Rk4Ep ep;
    };
    template <int L>
    static constexpr int lds()
    {
        return tt_lds<L, WG, typename GroupTTPlan<L>::P>();
    }
    template <int L>
    __device__ static void run(const Blk &bk, double2 *lds, const Args &a)
    {
        y_tt_update_h_body<L, WG, typename GroupTTPlan<L>::P, double2, EP>(
            bk, lds, a.W, a.M, a.f, a.result, a.N, a.Next, a.Ntot, a.kg,
            a.dscal, a.ls, a.tw, a.e, a.rw, a.ep);
    }
};

struct FamRK4Stage {
    static constexpr int WG = RED_WG;
    static constexpr bool BY_LEN = false;
    struct Args {
        double *out;
        const double *y, *k;
        double dt;
        int half, nt;
        long nl, tail;
    };
    template <int L> static constexpr int lds() { return 0; }
    template <int L>
    __device__ static void run(const Blk &bk, double2 *, const Args &a)
    {
        if (a.nt)
            rk4_stage_body<true>(bk.x, bk.gx, a.out, a.y, a.k, a.dt, a.half,
                                 a.nl, a.tail);
        else
            rk4_stage_body<false>(bk.x, bk.gx, a.out, a.y, a.k, a.dt, a.half,
                                  a.nl, a.tail);
    }
};

struct FamRK4Final {
    static constexpr int WG = RED_WG;
    static constexpr bool BY_LEN = false;
    struct Args {
        double *y;
        const double *k1, *k2, *k3, *k4;
        double dt;
        int nt;
        long nl, tail;
    };
    template <int L> static constexpr int lds() { return 0; }
    template <int L>
    __device__ static void run(const Blk &bk, double2 *, const Args &a)
    {
        if (a.nt)
            rk4_final_body<true>(bk.x, bk.gx, a.y, a.k1, a.k2, a.k3, a.k4,
                                 a.dt, a.nl, a.tail);
        else
            rk4_final_body<false>(bk.x, bk.gx, a.y, a.k1, a.k2, a.k3, a.k4,
                                  a.dt, a.nl, a.tail);
    }
};

// one segment of a grouped launch, before the table is laid out
template <class Fam>
struct Seg {
    int L;
    dim3 grid;
    typename Fam::Args a;
};

} // namespace ihg

struct ihg_group {
    std::vector<ihg_ctx *> c; // the contexts, in the caller's order
    std::vector<int> order;   // segment order: larger grids first
    hipStream_t stream;       // the first context's stream
    long launches = 0;        // grouped launches so far
};

namespace ihg {

// lay the segments out (in the group's order; segments of the same length
// keep the caller's order) and launch the family once
template <class Fam>
static int glaunch(ihg_group *g, const std::vector<Seg<Fam>> &segs,
                   const char *what)
{
    if (segs.empty()) return 0;
    if ((int)segs.size() > GMAX) return set_error("%s: too many segments", what);
    GroupTab<typename Fam::Args> t;
    std::memset(&t, 0, sizeof t);
    std::vector<int> idx(segs.size());
    for (size_t i = 0; i < segs.size(); ++i) idx[i] = (int)i;
    std::stable_sort(idx.begin(), idx.end(),
                     [&](int a, int b) { return segs[a].L > segs[b].L; });
    unsigned long begin = 0;
    int lds = 0, set = 0;
    t.n = (int)segs.size();
    for (int i = 0; i < t.n; ++i) {
        const Seg<Fam> &s = segs[idx[i]];
        lds = std::max(lds, fam_lds<Fam>(s.L));
        set |= len_bit(s.L);
        const unsigned long nb = (unsigned long)s.grid.x * s.grid.y * s.grid.z;
        if (nb == 0) return set_error("%s: empty segment", what);
        t.g[i] = SegGeom{(unsigned)begin, s.grid.x, s.grid.y, s.grid.z, s.L};
        t.a[i] = s.a;
        begin += (nb + 7) / 8 * 8;
        if (begin > 0x7fffffffUL) return set_error("%s: grid too large", what);
    }
    // the largest LDS any length of the family asks for is allowed once per
    // process and kernel (above 64 KB the runtime wants it declared)
    static const hipError_t attr[2] = {
        hipFuncSetAttribute(reinterpret_cast<const void *>(&k_group<Fam, SET_32_128>),
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(fam_lds_max<Fam>() * sizeof(double2))),
        hipFuncSetAttribute(reinterpret_cast<const void *>(&k_group<Fam, SET_ALL>),
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(fam_lds_max<Fam>() * sizeof(double2)))};
    IHG_TRY(hip_check(attr[0], "hipFuncSetAttribute"));
    IHG_TRY(hip_check(attr[1], "hipFuncSetAttribute"));
    const size_t shm = (size_t)lds * sizeof(double2);
    if ((set & ~SET_32_128) == 0)
        hipLaunchKernelGGL((k_group<Fam, SET_32_128>), dim3((unsigned)begin),
                           dim3(Fam::WG), shm, g->stream, t);
    else
        hipLaunchKernelGGL((k_group<Fam, SET_ALL>), dim3((unsigned)begin),
                           dim3(Fam::WG), shm, g->stream, t);
    g->launches += 1;
    return hip_check(hipGetLastError(), what);
}

static Ens ens_of(const ihg_ctx *c)
{
    return c->nreal > 1 ? c->ens : Ens{};
}

static LineSet y_lines_of(const ihg_ctx *c)
{
    return LineSet{(long)c->Nx * c->Mz, c->Mz, (long)c->Ny * c->Mz, c->Mz};
}

static LineSet x_lines_of(const ihg_ctx *c)
{
    const long plane = (long)c->Ny * c->Mz;
    return LineSet{plane, plane, 0, plane};
}

template <int DIR>
static Seg<FamStrided<DIR>> strided_seg(const ihg_ctx *c, double2 *data,
                                        LineSet ls, int batch, long cbatch)
{
    const int L = c->Nx; // cubic grids
    const int B = dispatch_len(L, [&](auto Lc) {
        return StridedShape<decltype(Lc)::value>::B;
    });
    Seg<FamStrided<DIR>> s;
    s.L = L;
    s.grid = dim3((unsigned)((ls.nlines + B - 1) / B), batch, c->nreal);
    s.a = {data, data, ls, cbatch, c->tw, ens_of(c).w};
    return s;
}

// lines per work-group of a z pass and of a strided (fan-out) pass
static int tile_b(int L)
{
    return dispatch_len(L, [&](auto Lc) { return Tile<decltype(Lc)::value>::B; });
}

static int strided_b(int L)
{
    return dispatch_len(L, [&](auto Lc) {
        return StridedShape<decltype(Lc)::value>::B;
    });
}

static unsigned upd_blocks(long n)
{
    long b = (n + RED_WG * 4 - 1) / (RED_WG * 4);
    if (b < 1) b = 1;
    return (unsigned)(b > 65536 ? 65536 : b);
}

static unsigned rk4_blocks(long npairs)
{
    long b = (npairs + RED_WG - 1) / RED_WG;
    if (b < 1) b = 1;
    return (unsigned)(b > 262144 ? 262144 : b);
}

// RK4 vectors of a context (as ihg_rk4_stage / ihg_rk4_final step them)
static void rk4_extent_of(const ihg_ctx *c, long &nl, long &tail)
{
    if (c->nreal > 1) {
        nl = c->nreal * c->ens.f;
        tail = -1;
    } else {
        nl = c->p.enable_gw ? c->Ntot - 1 : 4 * c->N;
        tail = c->Ntot - 1;
    }
}

} // namespace ihg

extern "C" int ihg_group_create(ihg_ctx *const *ctxs, int n, ihg_group **out)
{
    if (!ctxs || !out) return set_error("NULL argument");
    *out = nullptr;
    if (n < 1 || n > GMAX)
        return set_error("ihg_group_create: 1 to %d contexts, got %d", GMAX, n);
    for (int i = 0; i < n; ++i) {
        const ihg_ctx *c = ctxs[i];
        if (!c) return set_error("ihg_group_create: context %d is NULL", i);
        for (int j = 0; j < i; ++j)
            if (ctxs[j] == c)
                return set_error("ihg_group_create: context %d repeated", i);
        if (c->Nx != c->Ny || c->Ny != c->Nz || !group_len_ok(c->Nx))
            return set_error("ihg_group_create: context %d is %dx%dx%d; grouped "
                             "launches serve cubic grids of 16, 32, 64 or 128 "
                             "points", i, c->Nx, c->Ny, c->Nz);
        if (c->p.enable_gw && c->p.gw_fp32)
            return set_error("ihg_group_create: context %d has the FP32 GW "
                             "source; grouped launches are FP64", i);
    }
    ihg_group *g = new ihg_group();
    g->c.assign(ctxs, ctxs + n);
    g->stream = ctxs[0]->stream;
    *out = g;
    return 0;
}

extern "C" int ihg_group_destroy(ihg_group *g)
{
    delete g;
    return 0;
}

extern "C" long ihg_group_launches(const ihg_group *g)
{
    return g ? g->launches : -1;
}

static int check_arrays(const ihg_group *g, double *const *a, const char *name)
{
    if (!a) return set_error("NULL %s", name);
    for (size_t i = 0; i < g->c.size(); ++i)
        if (!a[i]) return set_error("NULL %s[%zu]", name, i);
    return 0;
}

// mk_rhs of every context, with the RK4 epilogue eps[i] of context i when
// eps is given; the launches of the three epilogue families go by whether
// eps is given (all contexts or none)
template <bool EP>
static int group_rhs(ihg_group *g, double *const *f, double *const *result,
                     unsigned flags, const Rk4Ep *eps)
{
    const int n = (int)g->c.size();
    std::vector<int> kept(n);
    for (int i = 0; i < n; ++i) {
        if (f[i] == result[i]) return set_error("f[%d] and result[%d] must differ", i, i);
        kept[i] = (flags & IHG_FILTER) && g->c[i]->p.enable_filter;
    }

    // mk_ffts_and_filter's filter branch (src/toolbox.c:164-171), as
    // rhs_filter: z r2c, y, x round trip with the window, y, z c2r of the
    // four fields
    {
        std::vector<Seg<FamZR2C>> zf;
        std::vector<Seg<FamStrided<-1>>> yf;
        std::vector<Seg<FamFanout<1, true, FilterFac>>> xf;
        std::vector<Seg<FamStrided<+1>>> yb;
        std::vector<Seg<FamZC2R>> zb;
        for (int i = 0; i < n; ++i) {
            if (!kept[i]) continue;
            ihg_ctx *c = g->c[i];
            const Ens e = ens_of(c);
            const long N = c->N, M = c->M, nrows = (long)c->Nx * c->Ny;
            const int L = c->Nx, B = tile_b(L), nr = c->nreal;
            const dim3 gz((unsigned)((nrows / 2 + B - 1) / B), 4, nr);
            zf.push_back({L, gz, {f[i], c->W, nrows, N, M, c->tw, e.f, e.w}});
            yf.push_back(strided_seg<-1>(c, c->W, y_lines_of(c), 4, M));
            const LineSet ls = x_lines_of(c);
            const FilterFac fac{c->filter, (long)c->Ny * c->Mz, c->Nx, c->Ny,
                                c->Mz, (double)N, c->W + 4 * M,
                                nr > 1 ? nullptr : c->Wc,
                                nr > 1 ? nullptr : c->Wc + M, e.w};
            const int Bx = strided_b(L);
            Seg<FamFanout<1, true, FilterFac>> sx;
            sx.L = L;
            sx.grid = dim3((unsigned)((ls.nlines + Bx - 1) / Bx), 4, nr);
            sx.a = {c->W, c->W, 0L, ls, M, c->tw, fac, e.w};
            xf.push_back(sx);
            yb.push_back(strided_seg<+1>(c, c->W, y_lines_of(c), 4, M));
            zb.push_back({L, gz, {c->W, f[i], nrows, M, N, c->tw, e.w, e.f}});
        }
        IHG_TRY(glaunch(g, zf, "grouped z r2c (filter)"));
        IHG_TRY(glaunch(g, yf, "grouped y pass (filter)"));
        IHG_TRY(glaunch(g, xf, "grouped x round trip (filter)"));
        IHG_TRY(glaunch(g, yb, "grouped y pass (filter)"));
        IHG_TRY(glaunch(g, zb, "grouped z c2r (filter)"));
    }

    // rhs_core: phi_hat (unless the filter kept it), the x fan-out of the
    // four derivative spectra, y backward, z c2r fused with rho and p, the
    // means and update_phi_psi (src/toolbox.c:75-299)
    {
        std::vector<Seg<FamZR2C>> z1;
        std::vector<Seg<FamStrided<-1>>> y1;
        std::vector<Seg<FamFanout<4, true, DerivFac>>> fw;
        std::vector<Seg<FamFanout<4, false, DerivFac>>> fk;
        std::vector<Seg<FamStrided<+1>>> yb;
        std::vector<Seg<FamZC2RRho>> zr;
        std::vector<Seg<FamMeans<EP>>> mn;
        std::vector<Seg<FamUpdate<EP>>> up;
        for (int i = 0; i < n; ++i) {
            ihg_ctx *c = g->c[i];
            const Ens e = ens_of(c);
            const long N = c->N, M = c->M, nrows = (long)c->Nx * c->Ny;
            const int L = c->Nx, B = tile_b(L), nr = c->nreal;
            double2 *W = c->W;
            if (!kept[i]) {
                const dim3 gz((unsigned)((nrows / 2 + B - 1) / B), 1, nr);
                z1.push_back({L, gz, {f[i], W + 4 * M, nrows, 0L, 0L, c->tw, e.f, e.w}});
                y1.push_back(strided_seg<-1>(c, W + 4 * M, y_lines_of(c), 1, 0L));
            }
            const LineSet ls = x_lines_of(c);
            const DerivFac fac{c->kg, (double)N};
            const int Bx = strided_b(L);
            const dim3 gx((unsigned)((ls.nlines + Bx - 1) / Bx), 1, nr);
            if (kept[i]) {
                Seg<FamFanout<4, false, DerivFac>> s;
                s.L = L;
                s.grid = gx;
                s.a = {W + 4 * M, W, M, ls, 0L, c->tw, fac, e.w};
                fk.push_back(s);
            } else {
                Seg<FamFanout<4, true, DerivFac>> s;
                s.L = L;
                s.grid = gx;
                s.a = {W + 4 * M, W, M, ls, 0L, c->tw, fac, e.w};
                fw.push_back(s);
            }
            yb.push_back(strided_seg<+1>(c, W, y_lines_of(c), 4, M));
            const unsigned nb = (unsigned)((2 * nrows + B - 1) / B);
            if (2L * nb > c->npartials)
                return set_error("partials buffer too small");
            const double m2 = c->p.mass * c->p.mass;
            Seg<FamZC2RRho> sr;
            sr.L = L;
            sr.grid = dim3(nb, 1, nr);
            sr.a = {W, M, f[i], result[i], c->p.enable_gw ? c->grad3 : nullptr,
                    c->partials, nrows, N, m2, c->tw, e};
            zr.push_back(sr);
            const Rk4Ep ep = EP ? eps[i] : Rk4Ep{};
            Seg<FamMeans<EP>> sm;
            sm.L = L;
            sm.grid = dim3(nr, 1, 1);
            sm.a = {c->partials, (int)nb, N, f[i], result[i], c->Ntot, c->dscal,
                    e, ep};
            mn.push_back(sm);
            Seg<FamUpdate<EP>> su;
            su.L = L;
            su.grid = dim3(upd_blocks(N), 1, nr);
            su.a = {f[i], result[i], N, c->Ntot, m2, c->dscal, e,
                    4.0 * 8.0 * (double)N * nr >= NT_BYTES ? 1 : 0, ep};
            up.push_back(su);
        }
        IHG_TRY(glaunch(g, z1, "grouped z r2c"));
        IHG_TRY(glaunch(g, y1, "grouped y pass"));
        IHG_TRY(glaunch(g, fw, "grouped x fan-out"));
        IHG_TRY(glaunch(g, fk, "grouped x fan-out (kept spectrum)"));
        IHG_TRY(glaunch(g, yb, "grouped y pass"));
        IHG_TRY(glaunch(g, zr, "grouped z c2r + rho, p"));
        IHG_TRY(glaunch(g, mn, "grouped means"));
        IHG_TRY(glaunch(g, up, "grouped update_phi_psi"));
    }

    // rhs_gw: z r2c of the six source products, x pass, the fused y pass +
    // TT projection + update_h (src/toolbox.c:311-424)
    {
        std::vector<Seg<FamSources>> sz;
        std::vector<Seg<FamStrided<-1>>> sx;
        std::vector<Seg<FamTT<EP>>> tt;
        for (int i = 0; i < n; ++i) {
            ihg_ctx *c = g->c[i];
            if (!c->p.enable_gw) continue;
            const Ens e = ens_of(c);
            const long N = c->N, M = c->M, nrows = (long)c->Nx * c->Ny;
            const int L = c->Nx, B = tile_b(L), nr = c->nreal;
            Seg<FamSources> s;
            s.L = L;
            s.grid = dim3((unsigned)((nrows + B - 1) / B), 1, nr);
            s.a = {c->grad3, c->W, M, nrows, N, c->tw, e.g, e.w};
            sz.push_back(s);
            sx.push_back(strided_seg<-1>(c, c->W, x_lines_of(c), 6, M));
            const LineSet ls = y_lines_of(c);
            const int Bt = dispatch_len(L, [&](auto Lc) {
                constexpr int LL = decltype(Lc)::value;
                return FamTT<EP>::WG / (LL / GroupTTPlan<LL>::P::E);
            });
            Seg<FamTT<EP>> st;
            st.L = L;
            st.grid = dim3((unsigned)((ls.nlines + Bt - 1) / Bt), 1, nr);
            st.a = {c->W, M, f[i], result[i], N, c->Next, c->Ntot, c->kg,
                    c->dscal, ls, c->tw, e, e.w, EP ? eps[i] : Rk4Ep{}};
            tt.push_back(st);
        }
        IHG_TRY(glaunch(g, sz, "grouped GW sources"));
        IHG_TRY(glaunch(g, sx, "grouped x pass (sources)"));
        IHG_TRY(glaunch(g, tt, "grouped y pass + TT + update_h"));
    }
    for (int i = 0; i < n; ++i) g->c[i]->calls_rhs += g->c[i]->nreal;
    return 0;
}

extern "C" int ihg_group_mk_rhs(ihg_group *g, double t, double *const *f,
                                double *const *result, unsigned flags)
{
    (void)t;
    if (!g) return set_error("NULL group");
    IHG_TRY(check_arrays(g, f, "f"));
    IHG_TRY(check_arrays(g, result, "result"));
    if (flags & IHG_OUTPUT)
        return set_error("ihg_group_mk_rhs: output calls are per realization");
    return group_rhs<false>(g, f, result, flags, nullptr);
}

extern "C" int ihg_group_mk_rhs_rk4(ihg_group *g, double t,
                                    double *const *f, double *const *result,
                                    unsigned flags, const ihg_rk4_ep *ep)
{
    (void)t;
    if (!g || !ep) return set_error("NULL argument");
    IHG_TRY(check_arrays(g, f, "f"));
    IHG_TRY(check_arrays(g, result, "result"));
    if (flags & IHG_OUTPUT)
        return set_error("ihg_group_mk_rhs_rk4: no output calls");
    const int n = (int)g->c.size();
    std::vector<Rk4Ep> eps(n);
    for (int i = 0; i < n; ++i) {
        const ihg_rk4_ep &x = ep[i];
        if (!x.y) return set_error("NULL ep[%d].y", i);
        Rk4Ep &d = eps[i];
        d = Rk4Ep{};
        d.dt = x.dt;
        d.y = x.y;
        if (x.mode != ep[0].mode)
            return set_error("ihg_group_mk_rhs_rk4: one mode for all contexts");
        if (x.mode == IHG_RK4_STAGE) {
            if (!x.out || x.out == f[i] || x.out == result[i] || x.out == x.y)
                return set_error("ihg_group_mk_rhs_rk4: ep[%d].out must be set "
                                 "and differ from f, result and y", i);
            d.mode = RK4_STAGE;
            d.half = x.half ? 1 : 0;
            d.out = x.out;
        } else if (x.mode == IHG_RK4_FINAL) {
            if (!x.k1 || !x.k2 || !x.k3)
                return set_error("NULL ep[%d].k1..k3", i);
            if (x.y == f[i] || x.y == result[i] || x.k1 == result[i] ||
                x.k2 == result[i] || x.k3 == result[i])
                return set_error("ihg_group_mk_rhs_rk4: ep[%d]: y and k1..k3 "
                                 "must differ from f and result", i);
            d.mode = RK4_FINAL;
            d.out = const_cast<double *>(x.y);
            d.k1 = x.k1;
            d.k2 = x.k2;
            d.k3 = x.k3;
        } else {
            return set_error("ihg_group_mk_rhs_rk4: mode %d", x.mode);
        }
    }
    return group_rhs<true>(g, f, result, flags, eps.data());
}

extern "C" int ihg_group_rk4_stage(ihg_group *g, double *const *out,
                                   double *const *y, double *const *k,
                                   double dt, int half)
{
    if (!g) return set_error("NULL group");
    IHG_TRY(check_arrays(g, out, "out"));
    IHG_TRY(check_arrays(g, y, "y"));
    IHG_TRY(check_arrays(g, k, "k"));
    std::vector<Seg<FamRK4Stage>> segs;
    for (size_t i = 0; i < g->c.size(); ++i) {
        const ihg_ctx *c = g->c[i];
        long nl, tail;
        rk4_extent_of(c, nl, tail);
        Seg<FamRK4Stage> s;
        s.L = c->Nx;
        s.grid = dim3(rk4_blocks(nl / 2), 1, 1);
        s.a = {out[i], y[i], k[i], dt, half ? 1 : 0,
               8.0 * (double)nl >= NT_BYTES ? 1 : 0, nl, tail};
        segs.push_back(s);
    }
    return glaunch(g, segs, "grouped RK4 stage");
}

extern "C" int ihg_group_rk4_final(ihg_group *g, double *const *y,
                                   double *const *k1, double *const *k2,
                                   double *const *k3, double *const *k4,
                                   double dt)
{
    if (!g) return set_error("NULL group");
    IHG_TRY(check_arrays(g, y, "y"));
    IHG_TRY(check_arrays(g, k1, "k1"));
    IHG_TRY(check_arrays(g, k2, "k2"));
    IHG_TRY(check_arrays(g, k3, "k3"));
    IHG_TRY(check_arrays(g, k4, "k4"));
    std::vector<Seg<FamRK4Final>> segs;
    for (size_t i = 0; i < g->c.size(); ++i) {
        const ihg_ctx *c = g->c[i];
        long nl, tail;
        rk4_extent_of(c, nl, tail);
        Seg<FamRK4Final> s;
        s.L = c->Nx;
        s.grid = dim3(rk4_blocks(nl / 2), 1, 1);
        s.a = {y[i], k1[i], k2[i], k3[i], k4[i], dt,
               8.0 * (double)nl >= NT_BYTES ? 1 : 0, nl, tail};
        segs.push_back(s);
    }
    return glaunch(g, segs, "grouped RK4 final");
}
