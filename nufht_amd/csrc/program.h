// Host-side description of an execution program (flat work lists for every
// stage over all boxes of one or many transforms), shared by the engine
// (engine.hip) and the asymptotic-stage launchers compiled per grid
// precision and exponent bound (asy_stage.hip, fft_stage.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <vector>

#include "device_common.h"
#include "fft_plan.h"
#include "internal.h"
#include "kernels_fused.h"

namespace nufht_dev {

// ---------------------------------------------------------------------------
// per-stage timing
// (one kernel per stage, except asy_fft: the rocFFT launches of all classes)
enum Stage { ST_ZERO, ST_PERM, ST_DIR, ST_LOCP, ST_LOCE, ST_GZERO, ST_SPREAD, ST_SPREAD_MF,
             ST_FFT, ST_INTERP, ST_INTERP_MF, ST_GRAD, ST_FUSED, ST_COUNT };
bool prof_enabled();
void prof_push(hipEvent_t a, hipEvent_t b, int stage);

struct StageTimer {
    int stage;
    hipStream_t st;
    hipEvent_t a = nullptr;
    StageTimer(int s, hipStream_t stream) : stage(s), st(stream)
    {
        if (prof_enabled()) {
            hipEventCreate(&a);
            hipEventRecord(a, st);
        }
    }
    ~StageTimer()
    {
        if (a) {
            hipEvent_t b;
            hipEventCreate(&b);
            hipEventRecord(b, st);
            prof_push(a, b, stage);
        }
    }
};

// sources per grid cell from which a box spreads on the matrix cores by
// default (NUFHT_AMD_DENSE_CELL overrides it)
double knobs_default_dense_cell();
// sources per VALU spread item before a position tile is split across items
constexpr int kSpreadSplit = 2048;
// consecutive sources per matrix-core spread item (0: from the plan's size)
constexpr int kDenseSplit = 0;
// sources and targets of an asymptotic box with a small uniform grid up to
// which it runs the fused kernel (kernels_fused.h); 0 by default: measured
// slower than the staged kernels on every benchmark workload, the autotuner
// tries it
constexpr int kFusedMax = 0;
int knobs_default_fused_max();   // kFusedMax, or NUFHT_AMD_FUSED_MAX

// Launch-geometry choices of a program (NufhtAmdKnobs without precision).
struct Knobs {
    double dense_cell = knobs_default_dense_cell();
    int dense_split = kDenseSplit;
    int spread_split = kSpreadSplit;
    int interp_mfma = 1;
    int own_fft = 1;
    int spread_sched = -1;   // matrix-core spread schedule: -1 automatic, 0 unified, 1 split
    int fused_max = knobs_default_fused_max();   // 0: no fused boxes
    bool operator==(const Knobs &o) const
    {
        return dense_cell == o.dense_cell && dense_split == o.dense_split &&
               spread_split == o.spread_split && interp_mfma == o.interp_mfma &&
               own_fft == o.own_fft && spread_sched == o.spread_sched &&
               fused_max == o.fused_max;
    }
    bool operator!=(const Knobs &o) const { return !(*this == o); }
};

struct FftClass {
    int nf = 0;
    long long batch = 0, off = 0;
    bool own = false;          // the two-pass FFT kernels of fft.h (else rocFFT, out of place)
    bool small = false;        // one pass (p1 only): lengths up to one LDS tile
    FftPass p1, p2;
    size_t lds1 = 0, lds2 = 0;
    rocfft_plan plan = nullptr;
    rocfft_execution_info info = nullptr;
    void *work = nullptr;
    size_t work_bytes = 0;
};

struct Program {
    Knobs knobs;      // set before build_program
    bool spread_split = false;   // matrix-core spread schedule (resolved from knobs)
    int prec = 0;     // 0 FP64, 2 mixed
    int emax = 0;
    std::vector<TfDesc> tfs;
    TfDesc *d_tfs = nullptr;
    std::vector<ZeroItem> zero;
    ZeroItem *d_zero = nullptr;
    std::vector<DirItem> dir;
    DirItem *d_dir = nullptr;
    std::vector<int2> dir_rng;        // source ranges {j0, count} of the direct items
    int dir_off[4] = {0, 0, 0, 0};    // direct items with <= 4, <= 8, <= 16 columns: [off[c], off[c+1])
    int2 *d_dir_rng = nullptr;
    std::vector<LocBoxD> lbox;
    LocBoxD *d_lbox = nullptr;
    std::vector<LocItem> lproj;
    LocItem *d_lproj = nullptr;
    std::vector<LocItem> leval1;      // targets covered by one local box
    LocItem *d_leval1 = nullptr;
    std::vector<LocItem> leval1h;     // the same for half-integer orders (no LDS scratch)
    LocItem *d_leval1h = nullptr;
    std::vector<int> lfold;           // half-integer local boxes, whose moments k_loc_fold folds
    int *d_lfold = nullptr;
    std::vector<LocEvalItem> leval;   // targets covered by several
    LocEvalItem *d_leval = nullptr;
    std::vector<int> lev_boxes;       // box lists of the evaluation items
    int *d_lev_boxes = nullptr;
    int loc_bmax = 16;                // columns per evaluation item (16 or 32)
    double *d_coef = nullptr;
    long long ncoef = 0;
    int loc_ch = 1024, loc_tpb = 256, loc_nlt = 1;
    size_t loc_lds_e = 0;
    std::vector<AsyBoxD> abox;
    std::vector<char> afused;         // box runs in the fused kernel
    // fused small boxes: units (in decreasing cost), their grid passes, the
    // work queue's counter, LDS grid elements per block, persistent blocks
    std::vector<FusedUnit> funits;
    std::vector<FftPass> fpass;
    FusedUnit *d_funits = nullptr;
    FftPass *d_fpass = nullptr;
    int *d_fqueue = nullptr;
    int fused_grid_elems = 0, fused_blocks = 0;
    AsyBoxD *d_abox = nullptr;
    std::vector<AsyItem> aspread, ainterp, azero, adense;
    std::vector<InterpRun> aimf;      // matrix-core interpolation runs
    std::vector<AsyItem> aimt;        // matrix-core interpolation tiles (knob interp_mfma 2)
    std::vector<int2> arange;
    std::vector<SpreadRange> adrange;
    std::vector<int> adense_off;      // matrix-core spread items by products per wave
    AsyItem *d_aspread = nullptr, *d_ainterp = nullptr, *d_azero = nullptr, *d_adense = nullptr;
    InterpRun *d_aimf = nullptr;
    AsyItem *d_aimt = nullptr;
    int2 *d_arange = nullptr;
    SpreadRange *d_adrange = nullptr;
    int interp_tmax = 16;
    void *d_grid = nullptr, *d_grid2 = nullptr;   // spread grid; transformed grid
    void *d_fftw = nullptr;                        // FFT twiddle tables
    std::vector<int4> fft_chinfo;                  // per channel of own FFT classes: L, mlo, span
    int4 *d_fft_chinfo = nullptr;
    long long grid_elems = 0;
    std::vector<FftClass> fft;
    double *d_ft = nullptr, *d_cheb = nullptr, *d_jtab = nullptr;
    long long jtab_total = 0;
    NufhtPlanStats stats;
    bool desc_dirty = true;
    // second stream of the near-field stages and its fork / join events
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
};

// Stage launchers of the asymptotic regime, instantiated per grid precision
// R (double, or float in mixed mode) and exponent bound EM (6, 10, 14, 22).
template <typename R> void asy_grid_zero(Program *P, hipStream_t st);
template <typename R> void asy_fft(Program *P, hipStream_t st);
template <typename R, int EM> void asy_spread(Program *P, hipStream_t st);
template <typename R, int EM> void asy_interp(Program *P, hipStream_t st);
template <typename R, int EM> void asy_fused(Program *P, hipStream_t st);

}  // namespace nufht_dev
