// plan.cpp -- host plan builders.  Compiled by the host C++ compiler with
// -ffp-contract=off: the integer tables come out of fp32 frequency arithmetic
// (powf, roundf) and must match the reference builder bit for bit.
#include "plan.hpp"

#include <algorithm>
#include <cmath>
#include <limits>

namespace sps {

int64_t checked_mul(int64_t a, int64_t b, const char* what) {
    if (a < 0 || b < 0 || (a != 0 && b > std::numeric_limits<int64_t>::max() / a))
        throw std::invalid_argument(std::string("size overflow: ") + what);
    return a * b;
}

int32_t EnginePlan::max_init_dt() const {
    int32_t m = 0;
    for (auto v : grid) m = std::max(m, v);
    return m;
}

namespace {

// Dispersion delay ratio between two bands (fdmtUtils.cpp:7-11): evaluated in
// fp32 with powf(f, -2).
float delay_ratio(float a_start, float a_end, float b_start, float b_end) {
    return (powf(a_start, -2.0f) - powf(a_end, -2.0f)) / (powf(b_start, -2.0f) - powf(b_end, -2.0f));
}

int64_t scaled_dt(float ratio, int64_t dt) { return (int64_t)roundf((float)dt * ratio); }

// Index of the grid entry nearest to val; on a tie the larger entry wins
// (fdmtUtils.cpp:63-78 uses lower_bound and steps left only when strictly
// closer).
int64_t nearest(const std::vector<int64_t>& g, int64_t val) {
    auto it = std::lower_bound(g.begin(), g.end(), val);
    if (it == g.end()) return (int64_t)g.size() - 1;
    int64_t i = it - g.begin();
    if (i > 0 && val - g[(size_t)i - 1] < g[(size_t)i] - val) --i;
    return i;
}

}  // namespace

// ---------------------------------------------------------------------------
FdmtPlan::FdmtPlan(float f_min_, float f_max_, int64_t nchans_, int64_t nsamps_, float tsamp_,
                   int64_t dt_max_, int64_t dt_step_, int64_t dt_min_)
    : f_min(f_min_), f_max(f_max_), tsamp(tsamp_), nchans(nchans_), nsamps(nsamps_), dt_max(dt_max_),
      dt_step(dt_step_), dt_min(dt_min_) {
    if (nchans < 1 || nsamps < 1) throw std::invalid_argument("nchans and nsamps must be positive");
    if (dt_step < 1) throw std::invalid_argument("dt_step must be positive");
    if (dt_min < 0 || dt_max < dt_min) throw std::invalid_argument("need 0 <= dt_min <= dt_max");
    if (!(f_max > f_min) || !(f_min > 0.f)) throw std::invalid_argument("need 0 < f_min < f_max");
    df = (f_max - f_min) / (float)nchans;
    correction = df / 2;
    niters = (int64_t)std::ceil(std::log2((double)nchans));
    iters.resize((size_t)niters + 1);

    // iteration 0: one sub-band per channel (fdmt_base.cpp:140-173)
    {
        IterPlan& it = iters[0];
        it.sub_row.push_back(0);
        for (int64_t c = 0; c < nchans; ++c) {
            float fs = df * (float)c + f_min;
            it.dt_grid.push_back(sub_grid(fs, fs + df));
            it.sub_row.push_back(it.sub_row.back() + (int64_t)it.dt_grid.back().size());
        }
        it.df_top = df;
        it.df_bot = df;
    }
    std::vector<int64_t> top_grid = iters[0].dt_grid.back();

    // merge iterations (fdmt_base.cpp:175-279)
    for (int64_t i = 1; i <= niters; ++i) {
        const IterPlan& prev = iters[(size_t)i - 1];
        IterPlan& it = iters[(size_t)i];
        const int64_t nprev = prev.nsub();
        const int64_t ncur = nprev / 2 + nprev % 2;
        const bool odd = (nprev % 2) == 1;
        it.df_top = odd ? prev.df_top : prev.df_top + prev.df_bot;
        it.df_bot = prev.df_bot * 2;
        it.sub_row.push_back(0);
        for (int64_t s = 0; s < ncur; ++s) {
            const int64_t tail_s = 2 * s, head_s = 2 * s + 1;
            const float fs = it.df_bot * (float)s + f_min;
            float fe, fm;
            std::vector<int64_t> grid;
            const bool top = (s == ncur - 1);
            if (top && odd) {
                fe = fs + it.df_top * 2;
                fm = fs + it.df_top;
                grid = top_grid;
            } else if (top) {
                fe = fs + it.df_top;
                fm = fs + it.df_bot / 2;
                grid = sub_grid(fs, fe);
                top_grid = grid;
            } else {
                fe = fs + it.df_bot;
                fm = fs + it.df_bot / 2;
                grid = sub_grid(fs, fe);
            }
            const float r_lo = delay_ratio(fs, fm - correction, fs, fe);
            const float r_hi = delay_ratio(fs, fm + correction, fs, fe);
            for (int64_t k = 0; k < (int64_t)grid.size(); ++k) {
                const int64_t dt = grid[(size_t)k];
                const int64_t mid_lo = scaled_dt(r_lo, dt);
                const int64_t mid_hi = scaled_dt(r_hi, dt);
                if (mid_lo > dt || mid_hi > dt) throw std::runtime_error("Invalid dt_mid values");
                Mapping m{};
                m.out = {s, k};
                if (top && odd) {
                    m.tail = {tail_s, nearest(prev.dt_grid[(size_t)tail_s], dt)};
                    m.head = {-1, -1};
                    m.offset = 0;
                    it.copies.push_back(m);
                } else {
                    if (mid_hi >= nsamps) throw std::runtime_error("Offset is greater than input size");
                    m.head = {head_s, nearest(prev.dt_grid[(size_t)head_s], dt - mid_hi)};
                    m.tail = {tail_s, nearest(prev.dt_grid[(size_t)tail_s], mid_lo)};
                    m.offset = mid_hi;
                    it.maps.push_back(m);
                }
            }
            it.dt_grid.push_back(std::move(grid));
            it.sub_row.push_back(it.sub_row.back() + (int64_t)it.dt_grid.back().size());
        }
    }
}

std::vector<int64_t> FdmtPlan::sub_grid(float f_start, float f_end) const {
    const float r = delay_ratio(f_start, f_end, f_min, f_max);
    const int64_t hi = scaled_dt(r, dt_max);
    const int64_t lo = scaled_dt(r, dt_min);
    std::vector<int64_t> g;
    for (int64_t dt = lo; dt <= hi; dt += dt_step) g.push_back(dt);
    return g;
}

void FdmtPlan::state_shape(int64_t i, int64_t out5[5]) const {
    const IterPlan& it = iters[(size_t)i];
    int64_t mn = std::numeric_limits<int64_t>::max(), mx = 0;
    for (const auto& g : it.dt_grid) {
        mn = std::min(mn, (int64_t)g.size());
        mx = std::max(mx, (int64_t)g.size());
    }
    out5[0] = it.nsub();
    out5[1] = mn;
    out5[2] = mx;
    out5[3] = it.nrows();
    out5[4] = nsamps;
}

std::vector<float> FdmtPlan::dm_grid_final() const {
    // fdmt_base.cpp:65-75 with kDispConst = 4.1493774e3 (fdmtUtils.hpp:7-11)
    const float k_disp = 4.1493774e3f;
    const float conv = k_disp * (powf(f_min, -2.0f) - powf(f_max, -2.0f));
    const float step = tsamp / conv;
    std::vector<float> out;
    for (auto dt : dt_grid_final()) out.push_back((float)dt * step);
    return out;
}

EnginePlan FdmtPlan::lower(int init_mode) const {
    if (init_mode != kInitMean && init_mode != kInitRestrict)
        throw std::invalid_argument("init mode must be mean or restrict for the plan-based FDMT");
    if (init_mode == kInitRestrict && dt_step != 1)
        throw std::invalid_argument("the restrict init form needs dt_step == 1");
    EnginePlan e;
    e.nchans = nchans;
    e.nsamps = nsamps;
    e.init_mode = init_mode;
    e.niters = (int)niters;
    const IterPlan& i0 = iters[0];
    e.grid_off.push_back(0);
    for (int64_t c = 0; c < nchans; ++c) {
        for (auto v : i0.dt_grid[(size_t)c]) {
            if (v > std::numeric_limits<int32_t>::max()) throw std::invalid_argument("dt too large");
            e.grid.push_back((int32_t)v);
        }
        e.grid_off.push_back((int64_t)e.grid.size());
    }
    e.chan_row_off = i0.sub_row;
    for (int64_t i = 0; i <= niters; ++i) e.rows.push_back(iters[(size_t)i].nrows());
    e.iter.resize((size_t)niters + 1);
    for (int64_t i = 1; i <= niters; ++i) {
        const IterPlan& it = iters[(size_t)i];
        const IterPlan& pv = iters[(size_t)i - 1];
        auto& rd = e.iter[(size_t)i];
        rd.reserve(it.maps.size() + it.copies.size());
        for (const auto& m : it.maps)
            rd.push_back({it.sub_row[(size_t)m.out.sub] + m.out.dt, pv.sub_row[(size_t)m.tail.sub] + m.tail.dt,
                          pv.sub_row[(size_t)m.head.sub] + m.head.dt, m.offset});
        for (const auto& m : it.copies)
            rd.push_back({it.sub_row[(size_t)m.out.sub] + m.out.dt, pv.sub_row[(size_t)m.tail.sub] + m.tail.dt, -1, 0});
        std::sort(rd.begin(), rd.end(), [](const RowDesc& a, const RowDesc& b) { return a.out_row < b.out_row; });
    }
    e.out_rows = e.rows.back();
    return e;
}

// ---------------------------------------------------------------------------
// legacy CFdmt
// ---------------------------------------------------------------------------
int CFdmtPlan::delta_t(float f0, float f1) const {
    // FdmtB.cpp:273-276: the squares are fp32, the rest double
    return (int)(std::ceil((1.0 / (f0 * f0) - 1.0 / (f1 * f1)) / (1.0 / (f_min * f_min) - 1.0 / (f_max * f_max)) *
                           (imaxdt - 1.0)));
}

int CFdmtPlan::init_delta() const { return delta_t(f_min, f_min + (f_max - f_min) / nchan); }

CFdmtPlan::CFdmtPlan(float f_min_, float f_max_, int nchan_, int cols_, int imaxdt_)
    : f_min(f_min_), f_max(f_max_), nchan(nchan_), cols(cols_), imaxdt(imaxdt_) {
    if (nchan < 1 || cols < 1 || imaxdt < 1) throw std::invalid_argument("nchan, cols and imaxDt must be positive");
    if (!(f_max > f_min) || !(f_min > 0.f)) throw std::invalid_argument("need 0 < Fmin < Fmax");
    // sub-matrix counts per iteration (FdmtB.cpp:213-270)
    quant.push_back(nchan);
    for (int q = nchan; q != 1;) {
        q = (q % 2 == 1) ? q / 2 + 1 : q / 2;
        quant.push_back(q);
        if (quant.size() > 100) throw std::runtime_error("too many iterations");
    }
    niter = (int)quant.size() - 1;
    // iteration 0 (FdmtCpu.cpp:166-179)
    std::vector<int> rows((size_t)nchan, init_delta() + 1);
    std::vector<float> f0((size_t)nchan + 1);
    for (int i = 0; i < nchan; ++i) f0[(size_t)i] = f_min + i * (f_max - f_min) / nchan;
    f0[(size_t)nchan] = f_max;
    freq.push_back(f0);
    auto csum = [](const std::vector<int>& r, int q) {
        std::vector<int> c((size_t)q + 1, 0);
        for (int i = 1; i <= q; ++i) c[(size_t)i] = c[(size_t)i - 1] + r[(size_t)i - 1];
        return c;
    };
    cumsum.push_back(csum(rows, nchan));
    // later iterations (calcNextStateConfig, FdmtB.cpp:278-299)
    for (int k = 1; k <= niter; ++k) {
        const int Q = quant[(size_t)k - 1];
        const std::vector<float>& F = freq[(size_t)k - 1];
        const int qc = quant[(size_t)k];
        std::vector<float> Fc((size_t)qc + 1);
        std::vector<int> rc((size_t)qc);
        for (int i = 0; i < Q / 2; ++i) {
            rc[(size_t)i] = delta_t(F[(size_t)2 * i], F[(size_t)2 * i + 2]) + 1;
            Fc[(size_t)i] = F[(size_t)2 * i];
        }
        if (Q % 2 == 1) {
            Fc[(size_t)qc - 1] = F[(size_t)Q - 1];
            rc[(size_t)qc - 1] = rows[(size_t)Q - 1];
        }
        Fc[(size_t)qc] = f_max;
        rows = rc;
        freq.push_back(Fc);
        cumsum.push_back(csum(rc, qc));
    }
}

EnginePlan CFdmtPlan::lower() const {
    EnginePlan e;
    e.nchans = nchan;
    e.nsamps = cols;
    e.init_mode = kInitCFdmt;
    e.niters = niter;
    const int D = init_delta();
    if (D < 0) throw std::runtime_error("negative initial delay");
    e.grid_off.push_back(0);
    e.chan_row_off.push_back(0);
    for (int c = 0; c < nchan; ++c) {
        for (int d = 0; d <= D; ++d) e.grid.push_back(d);
        e.grid_off.push_back((int64_t)e.grid.size());
        e.chan_row_off.push_back(e.chan_row_off.back() + D + 1);
    }
    // FdmtCpu.cpp:207-218 sizes state 0 from cumsum[0]; both equal nchan*(D+1)
    for (int k = 0; k <= niter; ++k) e.rows.push_back(cumsum[(size_t)k][(size_t)quant[(size_t)k]]);
    e.iter.resize((size_t)niter + 1);
    auto delay = [](float a, float b) { return (float)(1.0 / (a * a) - 1.0 / (b * b)); };  // FdmtCpu.cpp:369-372
    for (int k = 1; k <= niter; ++k) {
        const int Qp = quant[(size_t)k - 1];
        const auto& cs = cumsum[(size_t)k - 1];
        const auto& F = freq[(size_t)k - 1];
        const int Qc = quant[(size_t)k];
        const auto& csc = cumsum[(size_t)k];
        auto& rd = e.iter[(size_t)k];
        const int64_t prev_rows = e.rows[(size_t)k - 1];
        for (int i = 0; i < Qc; ++i) {
            if (i * 2 + 1 >= Qp) {
                // odd count: the last sub-matrix is copied (FdmtCpu.cpp:274-283)
                const int n = cs[(size_t)Qp] - cs[(size_t)Qp - 1];
                for (int j = 0; j < n; ++j) rd.push_back({(int64_t)csc[(size_t)i] + j, (int64_t)cs[(size_t)2 * i] + j, -1, 0});
                break;
            }
            const int n = csc[(size_t)i + 1] - csc[(size_t)i];
            const float c0 = delay(F[(size_t)2 * i], F[(size_t)2 * i + 1]) / delay(F[(size_t)2 * i], F[(size_t)2 * i + 2]);
            const float c1 = delay(F[(size_t)2 * i + 1], F[(size_t)2 * i + 2]) / delay(F[(size_t)2 * i], F[(size_t)2 * i + 2]);
            for (int j = 0; j < n; ++j) {
                const int j0 = (int)(c0 * ((float)j));
                const int j1 = (int)(c1 * ((float)j));
                RowDesc r{(int64_t)csc[(size_t)i] + j, (int64_t)cs[(size_t)2 * i] + j0, (int64_t)cs[(size_t)2 * i + 1] + j1,
                          (int64_t)j0};
                if (r.tail_row < 0 || r.tail_row >= prev_rows || r.head_row < 0 || r.head_row >= prev_rows)
                    throw std::runtime_error("legacy merge reads outside the previous state");
                rd.push_back(r);
            }
        }
    }
    e.out_rows = imaxdt;
    if (e.rows.back() < imaxdt) throw std::runtime_error("final state has fewer rows than imaxDt");
    return e;
}

}  // namespace sps
