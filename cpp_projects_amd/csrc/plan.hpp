// plan.hpp -- host-side FDMT plans and the flat row tables the HIP kernels run.
//
// Two plan flavours feed one execution engine:
//   * FdmtPlan   -- the plan-based FDMT (FDMT_PRAVIR_RESTRICT/fdmt_base.cpp:87-279)
//   * CFdmtPlan  -- the legacy CFdmt configuration (FDMT/FdmtB.cpp:158-299,
//                   FDMT/FdmtCpu.cpp:143-194,252-311)
// Both lower to an EnginePlan: per-channel init dt grids, then for every merge
// iteration a table of output rows {out_row, tail_row, head_row, offset}
// (head_row < 0 marks a copied row).  All indices are 64-bit: a 4096-channel
// x 2^18-sample state holds more than 2^31 elements.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace sps {

enum InitMode : int {
    kInitMean = 0,      // FDMT_PRAVIR running mean (fdmt_cpu.cpp:115-148)
    kInitRestrict = 1,  // FDMT_PRAVIR_RESTRICT prev*((dp+1)+sum)/(dc+1) (fdmt_cpu.cpp:91-114)
    kInitCFdmt = 2,     // CFdmtCpu fnc_initC (FdmtCpu.cpp:314-367)
};

struct RowDesc {
    int64_t out_row;
    int64_t tail_row;
    int64_t head_row;  // < 0: copy of tail_row
    int64_t offset;    // head is read at t - offset
};

struct EnginePlan {
    int64_t nchans = 0;
    int64_t nsamps = 0;
    int init_mode = kInitMean;
    // init: channel c owns rows [chan_row_off[c], chan_row_off[c+1]) of state 0
    // with dt values grid[grid_off[c] .. grid_off[c+1])
    std::vector<int32_t> grid;
    std::vector<int64_t> grid_off;
    std::vector<int64_t> chan_row_off;
    int niters = 0;
    std::vector<int64_t> rows;               // rows of state i, i = 0..niters
    std::vector<std::vector<RowDesc>> iter;  // iter[i] for i = 1..niters (iter[0] empty)
    int64_t out_rows = 0;                    // rows delivered to the caller
    int32_t max_init_dt() const;
};

// ---------------------------------------------------------------------------
// plan-based FDMT
// ---------------------------------------------------------------------------
struct Coord {
    int64_t sub, dt;
};
struct Mapping {
    Coord out, head, tail;
    int64_t offset;
};
struct IterPlan {
    float df_top = 0.f, df_bot = 0.f;
    std::vector<std::vector<int64_t>> dt_grid;  // per sub-band
    std::vector<int64_t> sub_row;               // first row of each sub-band (nsub + 1)
    std::vector<Mapping> maps;                  // summed coordinates
    std::vector<Mapping> copies;                // copied coordinates (head unused)
    int64_t nsub() const { return (int64_t)dt_grid.size(); }
    int64_t nrows() const { return sub_row.empty() ? 0 : sub_row.back(); }
};

class FdmtPlan {
public:
    FdmtPlan(float f_min, float f_max, int64_t nchans, int64_t nsamps, float tsamp, int64_t dt_max,
             int64_t dt_step, int64_t dt_min);
    float f_min, f_max, tsamp;
    int64_t nchans, nsamps, dt_max, dt_step, dt_min;
    float df, correction;
    int64_t niters;
    std::vector<IterPlan> iters;  // niters + 1
    // {nchans, ndt_min, ndt_max, sum ndt, nsamps} (fdmt_base.cpp:164-170)
    void state_shape(int64_t i, int64_t out5[5]) const;
    const std::vector<int64_t>& dt_grid_final() const { return iters[(size_t)niters].dt_grid[0]; }
    std::vector<float> dm_grid_final() const;
    EnginePlan lower(int init_mode) const;

private:
    std::vector<int64_t> sub_grid(float f_start, float f_end) const;
};

// ---------------------------------------------------------------------------
// legacy CFdmt
// ---------------------------------------------------------------------------
class CFdmtPlan {
public:
    CFdmtPlan(float f_min, float f_max, int nchan, int cols, int imaxdt);
    float f_min, f_max;
    int nchan, cols, imaxdt;
    int niter;
    std::vector<int> quant;                   // sub-matrices per iteration
    std::vector<std::vector<float>> freq;     // quant+1 edges per iteration
    std::vector<std::vector<int>> cumsum;     // quant+1 row offsets per iteration
    int init_delta() const;
    EnginePlan lower() const;

private:
    int delta_t(float f0, float f1) const;
};

// 64-bit cast helpers that refuse negative or overflowing products
int64_t checked_mul(int64_t a, int64_t b, const char* what);

}  // namespace sps
