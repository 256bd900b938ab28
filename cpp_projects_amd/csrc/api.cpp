// api.cpp -- the extern "C" boundary of libsps (declared in include/sps.h).
// Every entry point converts C++ exceptions into an sps_status and records
// the message for sps_last_error().
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "abi.hpp"
#include "common.hpp"
#include "kernels.hpp"
#include "plan.hpp"

using namespace sps;

namespace sps {
thread_local std::string g_last_error;
}

namespace {

template <class T>
void upload(DevBuf& b, const std::vector<T>& v, hipStream_t st) {
    b.ensure(std::max<size_t>(sizeof(T) * v.size(), 16));
    if (!v.empty()) SPS_HIP(hipMemcpyAsync(b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, st));
}

int bitlen(uint64_t v) {
    int b = 0;
    while (v) {
        ++b;
        v >>= 1;
    }
    return b;
}

int64_t round_up(int64_t a, int64_t m) { return ((a + m - 1) / m) * m; }

}  // namespace

// ---------------------------------------------------------------------------
// FDMT handle
// ---------------------------------------------------------------------------
struct sps_fdmt {
    std::unique_ptr<FdmtPlan> plan;    // plan-based FDMT
    std::unique_ptr<CFdmtPlan> cplan;  // legacy CFdmt
    EnginePlan eng;
    int64_t pitch = 0;                 // internal state pitch (floats)
    bool uploaded = false;
    DevBuf d_grid, d_grid_off, d_chan_row_off;
    std::vector<std::unique_ptr<DevBuf>> d_rows;  // per iteration: RowDescD of its live output rows
    std::vector<int64_t> d_rows_n;                // per iteration: number of those rows
    DevBuf state_a, state_b;
    int64_t rows_a = 0, rows_b = 0;    // capacity of the ping-pong states in rows
    DevBuf host_in, host_out;          // device copies for the host-pointer calls
    int dmax = 0;                      // largest init dt of any channel
    int grid_consec = 0;               // every channel's init dts are 0, 1, ..., K-1 with K <= 5
    // adjoint tables
    bool adj_built = false;
    std::vector<std::unique_ptr<DevBuf>> csr_off, csr_row, csr_shift;

    // pass schedule: fused iterations per pass (1..kMaxFuse, or 1..kMaxGroupFuse
    // for staged passes), summing to niters
    std::vector<int> schedule;
    std::vector<std::unique_ptr<DevBuf>> d_trees;  // per pass: TreeDesc of its live output rows
    std::vector<int64_t> tree_rows;                // per pass: number of those rows
    bool trees_ready = false;
    // staged (LDS-window) passes
    bool staged = false;
    bool fuse_init = false;    // first staged pass builds state 0 itself (mean init)
    bool ran_fused0 = false;   // the last execution did (stage accounting)
    bool staged_last = false;  // the final pass is staged too
    // wide-window fast tiles on the producer/consumer kernel; off where other
    // streams' kernels share the GPU (its two 74 KiB blocks per CU crowd them
    // out: a 4-stream batch of 64 beams took 236 instead of 225 ms)
    bool pc_fast = true;
    int group_floats = 8192;   // LDS window per tile buffer (floats; two buffers per block)
    int group_rows = 16;       // output rows per group at most
    int group_tiles = 8;       // consecutive 512-sample tiles per block
    static constexpr int kInitGroupFloats = 4096;  // first pass with the init fused (16 KiB per block)
    struct StagedPass {
        DevBuf groups, leaves, rows, unit_leaf;
        int64_t ngroups = 0;
        int lds_floats = 0;
        int64_t max_shift = 0;  // largest GRowDesc::maxshift
        bool init = false;     // leaves are ChanWinD (init fused)
    };
    std::vector<std::unique_ptr<StagedPass>> d_staged;  // per pass
    std::unique_ptr<StagedPass> d_stage0_init;          // first pass with the init fused

    void init_layout() {
        pitch = round_up(eng.nsamps, 4);
        dmax = eng.max_init_dt();
        grid_consec = 1;
        for (int64_t c = 0; c < eng.nchans && grid_consec; ++c) {
            const int64_t g0 = eng.grid_off[(size_t)c], K = eng.grid_off[(size_t)c + 1] - g0;
            if (K < 1 || K > 5) grid_consec = 0;
            for (int64_t k = 0; k < K && grid_consec; ++k)
                if (eng.grid[(size_t)(g0 + k)] != k) grid_consec = 0;
        }
        // state 0 in A; pass outputs alternate B, A, B, ... and rows only
        // shrink with the iteration, so A needs rows[0] and B rows[1]
        rows_a = eng.rows[0];
        rows_b = eng.niters >= 1 ? eng.rows[1] : 0;
        for (int i = 1; i <= eng.niters; ++i) {
            SPS_REQUIRE((int64_t)eng.iter[(size_t)i].size() == eng.rows[(size_t)i], "plan rows do not cover the state");
            for (int64_t r = 0; r < eng.rows[(size_t)i]; ++r)
                SPS_REQUIRE(eng.iter[(size_t)i][(size_t)r].out_row == r, "plan rows out of order");
            rows_a = std::max(rows_a, eng.rows[(size_t)i]);
        }
        compute_live();
        set_schedule(default_schedule(eng.niters));
    }

    // Rows of an intermediate state that no delivered output row depends on
    // (the plan's dt grid per sub-band is wider than what the last merges
    // read) are never computed: skipping them changes no output value.
    std::vector<std::vector<uint8_t>> live;  // per state
    DevBuf d_live0;                          // state-0 flags for the init kernels
    void compute_live() {
        const int n = eng.niters;
        live.assign((size_t)n + 1, {});
        for (int i = 0; i <= n; ++i) live[(size_t)i].assign((size_t)eng.rows[(size_t)i], 0);
        for (int64_t r = 0; r < eng.rows[(size_t)n]; ++r) live[(size_t)n][(size_t)r] = r < eng.out_rows ? 1 : 0;
        for (int i = n; i >= 1; --i)
            for (const auto& rd : eng.iter[(size_t)i]) {
                if (!live[(size_t)i][(size_t)rd.out_row]) continue;
                live[(size_t)i - 1][(size_t)rd.tail_row] = 1;
                if (rd.head_row >= 0) live[(size_t)i - 1][(size_t)rd.head_row] = 1;
            }
    }
    int64_t live_rows(int i) const {
        int64_t c = 0;
        for (auto v : live[(size_t)i]) c += v;
        return c;
    }

    static std::vector<int> default_schedule(int n) {
        std::vector<int> s;
        while (n > 0) {
            const int d = (n % kMaxFuse == 1 && n > kMaxFuse) ? 2 : std::min(n, kMaxFuse);
            s.push_back(d);
            n -= d;
        }
        return s;
    }

    void set_schedule(const std::vector<int>& s) {
        int sum = 0;
        const int maxd = staged ? kMaxGroupFuse : kMaxFuse;
        for (int d : s) {
            SPS_REQUIRE(d >= 1 && d <= maxd, staged ? "staged fusion depth must be 1..4" : "fusion depth must be 1..3");
            sum += d;
        }
        SPS_REQUIRE(sum == eng.niters, "schedule must cover every iteration");
        schedule = s;
        trees_ready = false;
    }

    // LDS window per tile buffer (floats), output rows per group and tiles
    // per block of staged passes
    void set_group_shape(int floats, int rows, int tiles) {
        SPS_REQUIRE(floats >= 2048 && floats <= 4 * kGroupMaxUnits, "group window must be 2048..9216 floats");
        SPS_REQUIRE(rows >= 1 && rows <= 64, "group rows must be 1..64");
        SPS_REQUIRE(tiles >= 1 && tiles <= 512, "tiles per block must be 1..512");
        group_floats = floats;
        group_rows = rows;
        group_tiles = tiles;
        trees_ready = false;
    }

    // 0: tree passes; 1: staged passes, the final pass on the tree kernel when
    // its depth allows; 2: as 1 with the init fused into the first pass (mean
    // init and at least two passes; otherwise as 1); 3, 4: as 1, 2 with the
    // final pass staged too
    void set_staged(int mode) {
        SPS_REQUIRE(mode >= 0 && mode <= 4, "staging mode must be 0..4");
        const bool on = mode > 0;
        const bool fi = mode == 2 || mode == 4, sl = mode >= 3;
        if (fuse_init != fi || staged_last != sl) trees_ready = false;
        fuse_init = fi;
        staged_last = sl;
        if (on == staged) return;
        staged = on;
        bool ok = true;
        for (int d : schedule) ok = ok && d <= (staged ? kMaxGroupFuse : kMaxFuse);
        if (!ok) schedule = default_schedule(eng.niters);
        trees_ready = false;
    }
    // the final pass of a staged schedule runs on the tree kernel
    bool tree_last() const { return staged && !staged_last && !schedule.empty() && schedule.back() <= kMaxFuse; }

    // Staged pass covering iterations i0 .. i0+L-1: group consecutive output
    // rows while their distinct leaves fit the LDS budget.
    // init_leaves: the pass reads the waterfall and builds its state-0 leaves
    // itself (first pass only); the LDS then holds every init row of each
    // channel the group touches.
    template <int L>
    bool build_groups_L(int i0, StagedPass& sp, hipStream_t st, bool init_leaves = false) const {
        const int top = i0 + L - 1;
        std::vector<int64_t> row_chan;
        if (init_leaves) {
            SPS_REQUIRE(i0 == 1, "only the first pass can build state 0");
            row_chan.resize((size_t)eng.rows[0]);
            for (int64_t c = 0; c < eng.nchans; ++c)
                for (int64_t q = eng.chan_row_off[(size_t)c]; q < eng.chan_row_off[(size_t)c + 1]; ++q)
                    row_chan[(size_t)q] = c;
        }
        auto chan_rows = [&](int64_t c) -> int64_t {
            return eng.chan_row_off[(size_t)c + 1] - eng.chan_row_off[(size_t)c];
        };
        const int64_t R = eng.rows[(size_t)top];
        const int NL = 1 << L;
        std::vector<int64_t> leaf((size_t)R * NL, -1), shift((size_t)R * NL, 0);
        std::vector<int32_t> offs((size_t)R * NL, -1);
        for (int64_t r = 0; r < R; ++r) {
            std::vector<int64_t> node((size_t)(2 << L), -1), sh((size_t)(2 << L), 0);
            node[1] = r;
            for (int k = 1; k < NL; ++k) {
                if (node[(size_t)k] < 0) continue;
                int dep = 0;
                while ((2 << dep) <= k) ++dep;
                const RowDesc& rd = eng.iter[(size_t)(top - dep)][(size_t)node[(size_t)k]];
                node[(size_t)(2 * k)] = rd.tail_row;
                sh[(size_t)(2 * k)] = sh[(size_t)k];
                if (rd.head_row >= 0) {
                    SPS_REQUIRE(rd.offset <= INT32_MAX, "offset too large");
                    offs[(size_t)r * NL + k] = (int32_t)rd.offset;
                    node[(size_t)(2 * k + 1)] = rd.head_row;
                    sh[(size_t)(2 * k + 1)] = sh[(size_t)k] + rd.offset;
                }
            }
            for (int j = 0; j < NL; ++j) {
                leaf[(size_t)r * NL + j] = node[(size_t)(NL + j)];
                shift[(size_t)r * NL + j] = sh[(size_t)(NL + j)];
            }
        }
        const int64_t tile = init_leaves ? kInitGroupTile : kGroupTile;
        auto units = [tile](int64_t smin, int64_t smax) -> int64_t { return (tile + round_up(smax, 4) - smin + 3) / 4; };
        constexpr size_t kMaxLeaves = 64;  // leaf indices of the unit table are bytes
        std::vector<uint8_t> unit_leaf;
        std::vector<GroupDescD> groups;
        std::vector<LeafDescD> leaves;
        std::vector<ChanWinD> chans;
        std::vector<GRowDesc<L>> rows;
        std::vector<int64_t> order;  // live output rows, in row order
        for (int64_t r = 0; r < R; ++r)
            if (live[(size_t)top][(size_t)r]) order.push_back(r);
        const int64_t nr = (int64_t)order.size();
        int maxf = 0;
        int64_t max_shift = 0;
        int64_t q = 0;
        while (q < nr) {
            std::map<int64_t, std::pair<int64_t, int64_t>> S;  // leaf row -> (smin, smax)
            const int64_t q0 = q;
            while (q < nr && q - q0 < group_rows) {
                const int64_t r = order[(size_t)q];
                auto S2 = S;
                for (int j = 0; j < NL; ++j) {
                    const int64_t lr = leaf[(size_t)r * NL + j];
                    if (lr < 0) continue;
                    const int64_t s = shift[(size_t)r * NL + j];
                    auto it = S2.find(lr);
                    if (it == S2.end())
                        S2.emplace(lr, std::make_pair(s, s));
                    else
                        it->second = {std::min(it->second.first, s), std::max(it->second.second, s)};
                }
                bool ok = S2.size() <= kMaxLeaves;
                int64_t floats = 0;
                if (init_leaves) {
                    // one window per channel, covering every use of any of its rows
                    std::map<int64_t, std::pair<int64_t, int64_t>> C;
                    for (const auto& kv : S2) {
                        const int64_t c = row_chan[(size_t)kv.first];
                        auto it = C.find(c);
                        if (it == C.end())
                            C.emplace(c, kv.second);
                        else
                            it->second = {std::min(it->second.first, kv.second.first),
                                          std::max(it->second.second, kv.second.second)};
                    }
                    ok = C.size() <= kMaxLeaves;
                    int64_t chunks = 0;
                    for (const auto& kv : C) {
                        const int64_t u = units(kv.second.first, kv.second.second);
                        floats += 4 * (u * chan_rows(kv.first) + 1);  // row 0 has a history chunk in front
                        chunks += u + 1;
                    }
                    ok = ok && chunks <= kInitGroupChunks;
                } else {
                    for (const auto& kv : S2) floats += 4 * units(kv.second.first, kv.second.second);
                }
                ok = ok && floats <= (init_leaves ? (int64_t)kInitGroupFloats : (int64_t)group_floats);
                if (!ok) {
                    if (q > q0) break;
                    // a lone tree over the budget: with init leaves the pass
                    // keeps the separate init; with state leaves it is taken
                    // anyway (2^L windows of at most 130 units each)
                    if (init_leaves) return false;
                }
                S.swap(S2);
                ++q;
            }
            GroupDescD g{};
            g.row0 = (int64_t)rows.size();
            g.nrows = (int32_t)(q - q0);
            std::map<int64_t, int32_t> base;
            int32_t off = 0;
            if (init_leaves) {
                std::map<int64_t, std::pair<int64_t, int64_t>> C;
                for (const auto& kv : S) {
                    const int64_t c = row_chan[(size_t)kv.first];
                    auto it = C.find(c);
                    if (it == C.end())
                        C.emplace(c, kv.second);
                    else
                        it->second = {std::min(it->second.first, kv.second.first),
                                      std::max(it->second.second, kv.second.second)};
                }
                g.leaf0 = (int64_t)chans.size();
                g.nleaves = (int32_t)C.size();
                g.unit0 = (int64_t)unit_leaf.size();
                int32_t xc = 0, ci = 0;
                for (const auto& kv : C) {
                    const int64_t c = kv.first;
                    ChanWinD cw{};
                    cw.chan = c;
                    cw.grid0 = eng.grid_off[(size_t)c];
                    cw.nrows = (int32_t)chan_rows(c);
                    cw.s4 = (int32_t)round_up(kv.second.second, 4);
                    cw.units = (int32_t)units(kv.second.first, kv.second.second);
                    cw.row_off = off;
                    cw.xc = xc;
                    // row 0 after its history chunk; rows k >= 1 after row 0
                    base[eng.chan_row_off[(size_t)c]] = off + 4 + cw.s4;
                    for (int64_t k = 1; k < cw.nrows; ++k)
                        base[eng.chan_row_off[(size_t)c] + k] =
                            off + 4 * (cw.units + 1) + (int32_t)(4 * cw.units * (k - 1)) + cw.s4;
                    off += 4 * (cw.units * cw.nrows + 1);
                    xc += cw.units + 1;
                    unit_leaf.insert(unit_leaf.end(), (size_t)cw.units + 1, (uint8_t)ci++);
                    chans.push_back(cw);
                }
                g.nunits = xc;
            } else {
                g.leaf0 = (int64_t)leaves.size();
                g.nleaves = (int32_t)S.size();
                g.unit0 = (int64_t)unit_leaf.size();
                int li = 0;
                for (const auto& kv : S) {
                    LeafDescD lf{};
                    lf.row = kv.first;
                    lf.smax4 = (int32_t)round_up(kv.second.second, 4);
                    lf.lds_off = off;
                    lf.nunits = (int32_t)units(kv.second.first, kv.second.second);
                    base[kv.first] = off + lf.smax4;
                    off += 4 * lf.nunits;
                    leaves.push_back(lf);
                    unit_leaf.insert(unit_leaf.end(), (size_t)lf.nunits, (uint8_t)li++);
                }
                g.nunits = off / 4;
                SPS_REQUIRE(g.nunits <= kGroupMaxUnits, "staged group window too large");
            }
            g.lds_floats = off;
            maxf = std::max(maxf, (int)off);
            groups.push_back(g);
            for (int64_t qq = q0; qq < q; ++qq) {
                const int64_t rr = order[(size_t)qq];
                GRowDesc<L> d{};
                d.out_row = rr;
                int64_t ms = 0;
                for (int j = 0; j < NL; ++j) {
                    const int64_t lr = leaf[(size_t)rr * NL + j];
                    d.lbase[j] = lr >= 0 ? base[lr] : 0;  // under a missing head: read, never added
                    if (lr >= 0) ms = std::max(ms, shift[(size_t)rr * NL + j]);
                }
                SPS_REQUIRE(ms <= INT32_MAX, "shift too large");
                d.maxshift = (int32_t)ms;
                max_shift = std::max(max_shift, ms);
                for (int k = 0; k < NL; ++k) d.off[k] = offs[(size_t)rr * NL + k];
                d.full = 1;
                for (int k = 1; k < NL; ++k)
                    if (d.off[k] < 0) d.full = 0;
                rows.push_back(d);
            }
        }
        upload(sp.groups, groups, st);
        if (init_leaves)
            upload(sp.leaves, chans, st);
        else
            upload(sp.leaves, leaves, st);
        upload(sp.rows, rows, st);
        upload(sp.unit_leaf, unit_leaf, st);
        sp.ngroups = (int64_t)groups.size();
        sp.lds_floats = maxf;
        sp.max_shift = max_shift;
        sp.init = init_leaves;
        return true;
    }

    template <int L>
    void build_trees_L(int i0, std::vector<uint8_t>& bytes) const {
        // pass covering iterations i0 .. i0+L-1 (input state i0-1)
        const int top = i0 + L - 1;
        std::vector<TreeDesc<L>> v;
        for (int64_t r = 0; r < eng.rows[(size_t)top]; ++r) {
            if (!live[(size_t)top][(size_t)r]) continue;  // dead rows are not computed
            v.emplace_back();
            TreeDesc<L>& d = v.back();
            d.out_row = r;
            for (auto& x : d.leaf) x = -1;
            for (auto& x : d.off) x = -1;
            // node k at depth dep corresponds to a row of iteration top - dep
            std::vector<int64_t> node((size_t)(2 << L), -1);
            node[1] = r;
            for (int k = 1; k < (1 << L); ++k) {
                if (node[(size_t)k] < 0) continue;
                int dep = 0;
                while ((2 << dep) <= k) ++dep;
                const RowDesc& rd = eng.iter[(size_t)(top - dep)][(size_t)node[(size_t)k]];
                node[(size_t)(2 * k)] = rd.tail_row;
                if (rd.head_row >= 0) {
                    SPS_REQUIRE(rd.offset <= INT32_MAX, "offset too large");
                    d.off[k] = (int32_t)rd.offset;
                    node[(size_t)(2 * k + 1)] = rd.head_row;
                }
            }
            for (int k = (1 << L); k < (2 << L); ++k) d.leaf[k - (1 << L)] = node[(size_t)k];
        }
        bytes.resize(sizeof(TreeDesc<L>) * v.size());
        std::memcpy(bytes.data(), v.data(), bytes.size());
        SPS_REQUIRE(!v.empty(), "a pass without live rows");
    }

    void ensure_trees(hipStream_t st) {
        if (trees_ready) return;
        d_trees.clear();
        tree_rows.clear();
        d_staged.clear();
        d_stage0_init.reset();
        if (staged) prepare_fdmt_group_kernels();
        int i0 = 1;
        const size_t np = schedule.size();
        d_trees.resize(np);
        tree_rows.assign(np, 0);
        d_staged.resize(np);
        for (size_t k = 0; k < np; ++k) {
            const int L = schedule[k];
            if (pass_staged(k)) {
                d_staged[k].reset(new StagedPass());
                StagedPass& sp = *d_staged[k];
                if (L == 1) build_groups_L<1>(i0, sp, st);
                else if (L == 2) build_groups_L<2>(i0, sp, st);
                else if (L == 3) build_groups_L<3>(i0, sp, st);
                else build_groups_L<4>(i0, sp, st);
            }
            // tree tables for every pass of depth <= 3: the tree passes use
            // them, and the mask transform runs on them in any mode
            if (L <= 3) {
                std::vector<uint8_t> bytes;
                if (L == 1) build_trees_L<1>(i0, bytes);
                else if (L == 2) build_trees_L<2>(i0, bytes);
                else build_trees_L<3>(i0, bytes);
                d_trees[k].reset(new DevBuf());
                upload(*d_trees[k], bytes, st);
                tree_rows[k] = live_rows(i0 + L - 1);
            }
            i0 += L;
        }
        if (staged) {
            if (fuse_init && eng.init_mode == kInitMean && grid_consec && np >= 2 && pass_staged(0) &&
                eng.nsamps <= kInitGroupMaxSamps) {
                // a first pass whose trees need more init rows than one LDS
                // window holds keeps the separate init
                std::unique_ptr<StagedPass> sp(new StagedPass());
                const int L = schedule[0];
                bool ok;
                if (L == 1) ok = build_groups_L<1>(1, *sp, st, true);
                else if (L == 2) ok = build_groups_L<2>(1, *sp, st, true);
                else if (L == 3) ok = build_groups_L<3>(1, *sp, st, true);
                else ok = build_groups_L<4>(1, *sp, st, true);
                if (ok) d_stage0_init = std::move(sp);
            }
        }
        SPS_HIP(hipStreamSynchronize(st));
        trees_ready = true;
    }
    bool pass_staged(size_t k) const { return staged && !(k + 1 == schedule.size() && tree_last()); }

    void ensure_device(hipStream_t st) {
        if (uploaded) return;
        upload(d_grid, eng.grid, st);
        upload(d_grid_off, eng.grid_off, st);
        upload(d_chan_row_off, eng.chan_row_off, st);
        d_rows.clear();
        d_rows_n.clear();
        for (int i = 0; i <= eng.niters; ++i) {
            d_rows.emplace_back(new DevBuf());
            std::vector<RowDescD> rd;
            for (const auto& r : eng.iter[(size_t)i])
                if (live[(size_t)i][(size_t)r.out_row]) rd.push_back({r.out_row, r.tail_row, r.head_row, r.offset});
            upload(*d_rows.back(), rd, st);
            d_rows_n.push_back((int64_t)rd.size());
        }
        upload(d_live0, live[0], st);
        SPS_HIP(hipStreamSynchronize(st));
        uploaded = true;
    }

    void ensure_states() {
        state_a.ensure(sizeof(float) * (size_t)checked_mul(std::max<int64_t>(rows_a, 1), pitch, "state"));
        if (eng.niters >= 2 || rows_b > 0)
            state_b.ensure(sizeof(float) * (size_t)checked_mul(std::max<int64_t>(rows_b, 1), pitch, "state"));
    }

    // Runs init + merges over nsamps samples (nsamps <= eng.nsamps) writing the
    // final rows to out (pitch out_pitch).  sa/sb: ping-pong states with pitch p.
    // ev (optional): one event recorded after the init and after every merge.
    // snr (optional): S/N epilogue fused into the last merge; aux: mask transform.
    void run(const void* in, int in_u8, int ones, int64_t in_pitch, const ChanNorm* norm, int64_t nsamps, float* out,
             int64_t out_pitch, float* sa, float* sb, int64_t p, hipStream_t st, hipEvent_t* ev = nullptr,
             const MergeEpilogue* snr = nullptr, bool aux = false) {
        SPS_REQUIRE(!snr || eng.niters > 0, "the S/N epilogue needs at least one merge");
        const ChanNorm* nm = norm;
        if (!aux) ran_fused0 = false;
        if (eng.niters == 0) {
            launch_fdmt_init(eng.init_mode, in, in_u8, ones, in_pitch, nm, eng.nchans, nsamps, d_grid.as<int32_t>(),
                             d_grid_off.as<int64_t>(), d_chan_row_off.as<int64_t>(), out, out_pitch, st, dmax,
                             nullptr, grid_consec);
            if (ev) SPS_HIP(hipEventRecord(ev[0], st));
            return;
        }
        ensure_trees(st);
        // the first staged pass builds state 0 itself when its tables exist
        const bool fused0 = !aux && !ones && in != nullptr && staged && d_stage0_init != nullptr;
        if (!aux) ran_fused0 = fused0;
        if (!fused0)
            launch_fdmt_init(eng.init_mode, in, in_u8, ones, in_pitch, nm, eng.nchans, nsamps, d_grid.as<int32_t>(),
                             d_grid_off.as<int64_t>(), d_chan_row_off.as<int64_t>(), sa, p, st, dmax,
                             d_live0.as<uint8_t>(), grid_consec);
        if (ev) SPS_HIP(hipEventRecord(ev[0], st));
        bool trees_all = true;
        for (int L : schedule) trees_all = trees_all && L <= 3;
        if (aux && trees_all) {
            // the mask transform over its short extent: fused tree passes
            // (the staged windows would stage mostly halo for so few tiles)
            int i0 = 1;
            for (size_t k = 0; k < schedule.size(); ++k) {
                const int L = schedule[k];
                const float* src = (k % 2 == 0) ? sa : sb;
                const bool last = (i0 + L - 1 == eng.niters);
                float* dst = last ? out : ((k % 2 == 0) ? sb : sa);
                launch_fdmt_fused(L, src, p, dst, last ? out_pitch : p, d_trees[k]->p, tree_rows[k], nsamps,
                                  last ? eng.out_rows : INT64_MAX, st, nullptr);
                i0 += L;
            }
            return;
        }
        if (aux) {
            // schedules with deeper passes: single-iteration launches
            for (int i = 1; i <= eng.niters; ++i) {
                const float* src = ((i - 1) % 2 == 0) ? sa : sb;
                const bool last = (i == eng.niters);
                float* dst = last ? out : ((i % 2 == 0) ? sa : sb);
                launch_fdmt_merge(src, p, dst, last ? out_pitch : p, d_rows[(size_t)i]->as<RowDescD>(),
                                  d_rows_n[(size_t)i], nsamps, last ? eng.out_rows : INT64_MAX, st, nullptr, true);
            }
            return;
        }
        int i0 = 1;
        for (size_t k = 0; k < schedule.size(); ++k) {
            const int L = schedule[k];
            const int top = i0 + L - 1;
            const float* src = (k % 2 == 0) ? sa : sb;
            const bool last = (top == eng.niters);
            float* dst = last ? out : ((k % 2 == 0) ? sb : sa);
            if (k == 0 && fused0) {
                const StagedPass& sp = *d_stage0_init;
                launch_fdmt_group_init(L, in, in_u8, in_pitch, nm, dst, p, sp.groups.p, sp.ngroups, sp.leaves.p,
                                       sp.rows.p, sp.unit_leaf.as<uint8_t>(), sp.lds_floats,
                                       group_tiles * (kGroupTile / kInitGroupTile), nsamps, sp.max_shift, st);
            } else if (pass_staged(k)) {
                const StagedPass& sp = *d_staged[k];
                launch_fdmt_group(L, src, p, dst, last ? out_pitch : p, sp.groups.p, sp.ngroups, sp.leaves.p,
                                  sp.rows.p, sp.unit_leaf.as<uint8_t>(), sp.lds_floats, group_tiles, nsamps,
                                  last ? eng.out_rows : INT64_MAX, sp.max_shift, st, last ? snr : nullptr, pc_fast);
            } else {
                launch_fdmt_fused(L, src, p, dst, last ? out_pitch : p, d_trees[k]->p, tree_rows[k], nsamps,
                                  last ? eng.out_rows : INT64_MAX, st, last ? snr : nullptr);
            }
            if (ev) SPS_HIP(hipEventRecord(ev[k + 1], st));
            i0 += L;
        }
    }

    void build_adjoint(hipStream_t st) {
        if (adj_built) return;
        csr_off.clear();
        csr_row.clear();
        csr_shift.clear();
        for (int i = 0; i <= eng.niters; ++i) {
            csr_off.emplace_back(new DevBuf());
            csr_row.emplace_back(new DevBuf());
            csr_shift.emplace_back(new DevBuf());
            if (i == 0) continue;
            const int64_t prev_rows = eng.rows[(size_t)i - 1];
            std::vector<std::vector<std::pair<int64_t, int64_t>>> users((size_t)prev_rows);
            for (const auto& r : eng.iter[(size_t)i]) {
                users[(size_t)r.tail_row].push_back({r.out_row, 0});
                if (r.head_row >= 0) users[(size_t)r.head_row].push_back({r.out_row, r.offset});
            }
            std::vector<int64_t> off(1, 0), row, sh;
            for (auto& u : users) {
                std::sort(u.begin(), u.end());
                for (auto& e : u) {
                    row.push_back(e.first);
                    sh.push_back(e.second);
                }
                off.push_back((int64_t)row.size());
            }
            upload(*csr_off.back(), off, st);
            upload(*csr_row.back(), row, st);
            upload(*csr_shift.back(), sh, st);
        }
        SPS_HIP(hipStreamSynchronize(st));
        adj_built = true;
    }
};

static void finish_create(sps_fdmt* h) {
    h->init_layout();
}

extern "C" {

SPS_API const char* sps_last_error(void) { return g_last_error.c_str(); }
SPS_API int sps_version(void) { return 100; }

SPS_API sps_status sps_device_count(int* h_count) {
    return guarded([&] {
        SPS_REQUIRE(h_count, "null output");
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        *h_count = (e == hipSuccess) ? n : 0;
    });
}

SPS_API sps_status sps_fdmt_create(float f_min, float f_max, int64_t nchans, int64_t nsamps, float tsamp,
                                   int64_t dt_max, int64_t dt_step, int64_t dt_min, int init_mode, sps_fdmt** out) {
    return guarded([&] {
        SPS_REQUIRE(out, "null output handle");
        *out = nullptr;
        std::unique_ptr<sps_fdmt> h(new sps_fdmt());
        h->plan.reset(new FdmtPlan(f_min, f_max, nchans, nsamps, tsamp, dt_max, dt_step, dt_min));
        h->eng = h->plan->lower(init_mode);
        finish_create(h.get());
        *out = h.release();
    });
}

SPS_API sps_status sps_cfdmt_create(float f_min, float f_max, int nchan, int cols, int imaxdt, sps_fdmt** out) {
    return guarded([&] {
        SPS_REQUIRE(out, "null output handle");
        *out = nullptr;
        std::unique_ptr<sps_fdmt> h(new sps_fdmt());
        h->cplan.reset(new CFdmtPlan(f_min, f_max, nchan, cols, imaxdt));
        h->eng = h->cplan->lower();
        finish_create(h.get());
        *out = h.release();
    });
}

SPS_API sps_status sps_fdmt_destroy(sps_fdmt* h) {
    return guarded([&] { delete h; });
}

SPS_API sps_status sps_fdmt_get_niters(const sps_fdmt* h, int64_t* n) {
    return guarded([&] {
        SPS_REQUIRE(h && n, "null argument");
        *n = h->eng.niters;
    });
}

SPS_API sps_status sps_fdmt_get_state_shape(const sps_fdmt* h, int64_t iter, int64_t* s5) {
    return guarded([&] {
        SPS_REQUIRE(h && s5, "null argument");
        SPS_REQUIRE(iter >= 0 && iter <= h->eng.niters, "iteration out of range");
        if (h->plan) {
            h->plan->state_shape(iter, s5);
        } else {
            const auto& c = *h->cplan;
            const int q = c.quant[(size_t)iter];
            int mn = INT32_MAX, mx = 0;
            for (int i = 0; i < q; ++i) {
                int r = c.cumsum[(size_t)iter][(size_t)i + 1] - c.cumsum[(size_t)iter][(size_t)i];
                mn = std::min(mn, r);
                mx = std::max(mx, r);
            }
            s5[0] = q;
            s5[1] = mn;
            s5[2] = mx;
            s5[3] = c.cumsum[(size_t)iter][(size_t)q];
            s5[4] = c.cols;
        }
    });
}

SPS_API sps_status sps_fdmt_get_output_shape(const sps_fdmt* h, int64_t* rows, int64_t* cols) {
    return guarded([&] {
        SPS_REQUIRE(h && rows && cols, "null argument");
        *rows = h->eng.out_rows;
        *cols = h->eng.nsamps;
    });
}

SPS_API sps_status sps_fdmt_get_dt_grid_final(const sps_fdmt* h, int64_t* outv, int64_t cap, int64_t* n) {
    return guarded([&] {
        SPS_REQUIRE(h && n, "null argument");
        std::vector<int64_t> g;
        if (h->plan) g = h->plan->dt_grid_final();
        else
            for (int i = 0; i < h->cplan->imaxdt; ++i) g.push_back(i);
        *n = (int64_t)g.size();
        if (outv) std::copy(g.begin(), g.begin() + std::min<int64_t>(cap, (int64_t)g.size()), outv);
    });
}

SPS_API sps_status sps_fdmt_get_dm_grid_final(const sps_fdmt* h, float* outv, int64_t cap, int64_t* n) {
    return guarded([&] {
        SPS_REQUIRE(h && n, "null argument");
        SPS_REQUIRE(h->plan, "DM grid is defined for plan-based FDMT handles");
        auto g = h->plan->dm_grid_final();
        *n = (int64_t)g.size();
        if (outv) std::copy(g.begin(), g.begin() + std::min<int64_t>(cap, (int64_t)g.size()), outv);
    });
}

SPS_API sps_status sps_fdmt_get_df(const sps_fdmt* h, float* df, float* corr) {
    return guarded([&] {
        SPS_REQUIRE(h && df && corr, "null argument");
        SPS_REQUIRE(h->plan, "df is defined for plan-based FDMT handles");
        *df = h->plan->df;
        *corr = h->plan->correction;
    });
}

SPS_API sps_status sps_fdmt_get_plan_iter(const sps_fdmt* h, int64_t iter, int64_t* nsub, int64_t* grid_off,
                                          int64_t* grid, int64_t* nmap, int64_t* map7, int64_t* ncopy,
                                          int64_t* copy4, float* df_top, float* df_bot) {
    return guarded([&] {
        SPS_REQUIRE(h, "null handle");
        SPS_REQUIRE(h->plan, "plan tables are defined for plan-based FDMT handles");
        SPS_REQUIRE(iter >= 0 && iter <= h->plan->niters, "iteration out of range");
        const IterPlan& it = h->plan->iters[(size_t)iter];
        if (nsub) *nsub = it.nsub();
        if (grid_off) {
            grid_off[0] = 0;
            for (int64_t s = 0; s < it.nsub(); ++s) grid_off[s + 1] = grid_off[s] + (int64_t)it.dt_grid[(size_t)s].size();
        }
        if (grid) {
            int64_t k = 0;
            for (const auto& g : it.dt_grid)
                for (auto v : g) grid[k++] = v;
        }
        if (nmap) *nmap = (int64_t)it.maps.size();
        if (map7)
            for (size_t k = 0; k < it.maps.size(); ++k) {
                const Mapping& m = it.maps[k];
                int64_t* r = map7 + 7 * k;
                r[0] = m.out.sub; r[1] = m.out.dt; r[2] = m.head.sub; r[3] = m.head.dt;
                r[4] = m.tail.sub; r[5] = m.tail.dt; r[6] = m.offset;
            }
        if (ncopy) *ncopy = (int64_t)it.copies.size();
        if (copy4)
            for (size_t k = 0; k < it.copies.size(); ++k) {
                const Mapping& m = it.copies[k];
                int64_t* r = copy4 + 4 * k;
                r[0] = m.out.sub; r[1] = m.out.dt; r[2] = m.tail.sub; r[3] = m.tail.dt;
            }
        if (df_top) *df_top = it.df_top;
        if (df_bot) *df_bot = it.df_bot;
    });
}

SPS_API sps_status sps_fdmt_get_rows(const sps_fdmt* h, int64_t iter, int64_t* n, int64_t* rows4) {
    return guarded([&] {
        SPS_REQUIRE(h && n, "null argument");
        SPS_REQUIRE(iter >= 1 && iter <= h->eng.niters, "iteration out of range");
        const auto& rd = h->eng.iter[(size_t)iter];
        *n = (int64_t)rd.size();
        if (rows4)
            for (size_t k = 0; k < rd.size(); ++k) {
                rows4[4 * k + 0] = rd[k].out_row;
                rows4[4 * k + 1] = rd[k].tail_row;
                rows4[4 * k + 2] = rd[k].head_row;
                rows4[4 * k + 3] = rd[k].offset;
            }
    });
}

SPS_API sps_status sps_fdmt_execute(sps_fdmt* h, const void* d_wf, int dtype, size_t wf_size, float* d_dmt,
                                    size_t dmt_size, int ones, void* stream) {
    return guarded([&] {
        SPS_REQUIRE(h, "null handle");
        const EnginePlan& e = h->eng;
        if (!ones) {
            SPS_REQUIRE(d_wf, "null waterfall");
            if ((int64_t)wf_size != e.nchans * e.nsamps) throw Error(SPS_ERR_INVALID_ARGUMENT, "Invalid size of waterfall");
        }
        SPS_REQUIRE(dtype == SPS_F32 || dtype == SPS_U8, "unknown dtype");
        SPS_REQUIRE(d_dmt, "null dmt");
        if ((int64_t)dmt_size != e.out_rows * e.nsamps) throw Error(SPS_ERR_INVALID_ARGUMENT, "Invalid size of dmt");
        hipStream_t st = as_stream(stream);
        h->ensure_device(st);
        h->ensure_states();
        h->run(d_wf, dtype == SPS_U8, ones, e.nsamps, nullptr, e.nsamps, d_dmt, e.nsamps, h->state_a.as<float>(),
               h->state_b.as<float>(), h->pitch, st);
    });
}

SPS_API sps_status sps_fdmt_initialise(sps_fdmt* h, const void* d_wf, int dtype, float* d_state, size_t state_size,
                                       void* stream) {
    return guarded([&] {
        SPS_REQUIRE(h && d_wf && d_state, "null argument");
        const EnginePlan& e = h->eng;
        SPS_REQUIRE((int64_t)state_size == e.rows[0] * e.nsamps, "Invalid size of state");
        hipStream_t st = as_stream(stream);
        h->ensure_device(st);
        launch_fdmt_init(e.init_mode, d_wf, dtype == SPS_U8, 0, e.nsamps, nullptr, e.nchans, e.nsamps,
                         h->d_grid.as<int32_t>(), h->d_grid_off.as<int64_t>(), h->d_chan_row_off.as<int64_t>(), d_state,
                         e.nsamps, st, h->dmax, nullptr, h->grid_consec);
    });
}

SPS_API sps_status sps_fdmt_execute_host(sps_fdmt* h, const float* h_waterfall, size_t waterfall_size, float* h_dmt,
                                         size_t dmt_size, int ones) {
    return guarded([&] {
        SPS_REQUIRE(h && h_dmt, "null argument");
        const EnginePlan& e = h->eng;
        if (!ones) {
            SPS_REQUIRE(h_waterfall, "null waterfall");
            if ((int64_t)waterfall_size != e.nchans * e.nsamps)
                throw Error(SPS_ERR_INVALID_ARGUMENT, "Invalid size of waterfall");
        }
        if ((int64_t)dmt_size != e.out_rows * e.nsamps) throw Error(SPS_ERR_INVALID_ARGUMENT, "Invalid size of dmt");
        hipStream_t st = nullptr;
        h->ensure_device(st);
        h->ensure_states();
        const void* d_in = nullptr;
        if (!ones) {
            h->host_in.ensure(sizeof(float) * waterfall_size);
            SPS_HIP(hipMemcpyAsync(h->host_in.p, h_waterfall, sizeof(float) * waterfall_size, hipMemcpyHostToDevice,
                                   st));
            d_in = h->host_in.p;
        }
        h->host_out.ensure(sizeof(float) * dmt_size);
        h->run(d_in, 0, ones, e.nsamps, nullptr, e.nsamps, h->host_out.as<float>(), e.nsamps, h->state_a.as<float>(),
               h->state_b.as<float>(), h->pitch, st);
        SPS_HIP(hipMemcpyAsync(h_dmt, h->host_out.p, sizeof(float) * dmt_size, hipMemcpyDeviceToHost, st));
        SPS_HIP(hipStreamSynchronize(st));
    });
}

SPS_API sps_status sps_fdmt_initialise_host(sps_fdmt* h, const float* h_waterfall, float* h_state, size_t state_size) {
    return guarded([&] {
        SPS_REQUIRE(h && h_waterfall && h_state, "null argument");
        const EnginePlan& e = h->eng;
        SPS_REQUIRE((int64_t)state_size == e.rows[0] * e.nsamps, "Invalid size of state");
        hipStream_t st = nullptr;
        h->ensure_device(st);
        const size_t nwf = (size_t)(e.nchans * e.nsamps);
        h->host_in.ensure(sizeof(float) * nwf);
        h->host_out.ensure(sizeof(float) * state_size);
        SPS_HIP(hipMemcpyAsync(h->host_in.p, h_waterfall, sizeof(float) * nwf, hipMemcpyHostToDevice, st));
        launch_fdmt_init(e.init_mode, h->host_in.p, 0, 0, e.nsamps, nullptr, e.nchans, e.nsamps,
                         h->d_grid.as<int32_t>(), h->d_grid_off.as<int64_t>(), h->d_chan_row_off.as<int64_t>(),
                         h->host_out.as<float>(), e.nsamps, st, h->dmax, nullptr, h->grid_consec);
        SPS_HIP(hipMemcpyAsync(h_state, h->host_out.p, sizeof(float) * state_size, hipMemcpyDeviceToHost, st));
        SPS_HIP(hipStreamSynchronize(st));
    });
}

SPS_API sps_status sps_cfdmt_get_config(const sps_fdmt* h, int64_t iter, int32_t* h_quant, int32_t* h_rows_cumsum,
                                        float* h_freq) {
    return guarded([&] {
        SPS_REQUIRE(h && h_quant, "null argument");
        SPS_REQUIRE(h->cplan, "not a legacy CFdmt handle");
        const CFdmtPlan& c = *h->cplan;
        SPS_REQUIRE(iter >= 0 && iter <= c.niter, "iteration out of range");
        const int q = c.quant[(size_t)iter];
        *h_quant = q;
        for (int k = 0; k <= q; ++k) {
            if (h_rows_cumsum) h_rows_cumsum[k] = c.cumsum[(size_t)iter][(size_t)k];
            if (h_freq) h_freq[k] = c.freq[(size_t)iter][(size_t)k];
        }
    });
}

SPS_API sps_status sps_fdmt_adjoint(sps_fdmt* h, const float* g_dmt, size_t dmt_size, float* g_wf, size_t wf_size,
                                    void* stream) {
    return guarded([&] {
        SPS_REQUIRE(h && g_dmt && g_wf, "null argument");
        const EnginePlan& e = h->eng;
        SPS_REQUIRE(e.init_mode != kInitRestrict, "the restrict init form is not linear; no adjoint");
        SPS_REQUIRE((int64_t)dmt_size == e.out_rows * e.nsamps, "Invalid size of dmt");
        SPS_REQUIRE((int64_t)wf_size == e.nchans * e.nsamps, "Invalid size of waterfall");
        SPS_REQUIRE(e.rows.back() == e.out_rows, "adjoint needs the full final state");
        hipStream_t st = as_stream(stream);
        h->ensure_device(st);
        h->ensure_states();
        h->build_adjoint(st);
        const int64_t N = e.nsamps, P = h->pitch;
        float* sa = h->state_a.as<float>();
        float* sb = h->state_b.as<float>();
        const float* cur = g_dmt;
        int64_t cur_pitch = N;
        for (int i = e.niters; i >= 1; --i) {
            float* dst = ((i - 1) % 2 == 0) ? sa : sb;
            launch_fdmt_merge_adjoint(cur, cur_pitch, dst, P, h->csr_off[(size_t)i]->as<int64_t>(),
                                      h->csr_row[(size_t)i]->as<int64_t>(), h->csr_shift[(size_t)i]->as<int64_t>(),
                                      e.rows[(size_t)i - 1], N, st);
            cur = dst;
            cur_pitch = P;
        }
        launch_fdmt_init_adjoint(cur, cur_pitch, g_wf, N, e.nchans, N, h->d_grid.as<int32_t>(),
                                 h->d_grid_off.as<int64_t>(), h->d_chan_row_off.as<int64_t>(), st);
    });
}

// ---------------------------------------------------------------------------
// normalisation
// ---------------------------------------------------------------------------
SPS_API sps_status sps_calc_fdmt_inputs(float* d_in, int64_t nrows, int64_t ncols, float* d_buff, float* d_norm_input,
                                        int mode, void* stream) {
    return guarded([&] {
        SPS_REQUIRE(d_in && nrows > 0 && ncols > 0, "invalid input");
        SPS_REQUIRE(mode == SPS_NORM_GLOBAL || mode == SPS_NORM_PER_CHANNEL, "mode must be global or per-channel");
        hipStream_t st = as_stream(stream);
        DevBuf rm, rv, stats, mu, sg, keep;
        rm.ensure(sizeof(double) * (size_t)nrows);
        rv.ensure(sizeof(double) * (size_t)nrows);
        stats.ensure(sizeof(double) * 8);
        mu.ensure(sizeof(float) * (size_t)nrows);
        sg.ensure(sizeof(float) * (size_t)nrows);
        keep.ensure((size_t)nrows);
        launch_row_stats(d_in, 0, nrows, ncols, rm.as<double>(), rv.as<double>(), st);
        launch_stats_finalize(rm.as<double>(), rv.as<double>(), nrows, mode, stats.as<double>(), mu.as<float>(),
                              sg.as<float>(), keep.as<uint8_t>(), d_buff, st);
        launch_apply_norm(d_in, d_norm_input, nrows, ncols, mu.as<float>(), sg.as<float>(), keep.as<uint8_t>(), st);
        SPS_HIP(hipStreamSynchronize(st));
    });
}

SPS_API sps_status sps_fdmt_normalization(const float* d_arr, const float* d_norm, int64_t len, float* d_out,
                                          void* stream) {
    return guarded([&] {
        SPS_REQUIRE(d_arr && d_norm && d_out && len >= 0, "invalid argument");
        launch_fdmt_normalization(d_arr, d_norm, len, d_out, as_stream(stream));
    });
}

SPS_API sps_status sps_fdmt_normalization_rows(const float* d_arr, int64_t rows, int64_t cols, const float* d_norm,
                                               int64_t norm_cols, float* d_out, void* stream) {
    return guarded([&] {
        SPS_REQUIRE(d_arr && d_norm && d_out && rows >= 0 && cols >= 0, "invalid argument");
        SPS_REQUIRE(norm_cols >= 1 && norm_cols <= cols, "norm_cols must lie in 1..cols");
        launch_fdmt_normalization_rows(d_arr, rows, cols, d_norm, norm_cols, d_out, as_stream(stream));
    });
}

// ---------------------------------------------------------------------------
// candidates
// ---------------------------------------------------------------------------
namespace {

// Sorts n candidates in place by (mdt, mt, mwidth).
void sort_cands(CandD* c, int64_t n, int64_t rows, int64_t cols, int64_t wmax, hipStream_t st) {
    if (n <= 1) return;
    const int bt = std::max(1, bitlen((uint64_t)cols)), bw = std::max(1, bitlen((uint64_t)wmax)),
              br = std::max(1, bitlen((uint64_t)rows));
    SPS_REQUIRE(bt + bw + br <= 64, "candidate key does not fit 64 bits");
    DevBuf k, kt, v, vt, hs;
    k.ensure(sizeof(uint64_t) * (size_t)n);
    kt.ensure(sizeof(uint64_t) * (size_t)n);
    v.ensure(sizeof(uint32_t) * (size_t)n);
    vt.ensure(sizeof(uint32_t) * (size_t)n);
    const int64_t hc = radix_hist_entries(n);
    hs.ensure(sizeof(uint32_t) * (size_t)hc);
    launch_cand_keys(c, n, bt, bw, k.as<uint64_t>(), v.as<uint32_t>(), st);
    radix_sort_pairs(k.as<uint64_t>(), v.as<uint32_t>(), kt.as<uint64_t>(), vt.as<uint32_t>(), n, bt + bw + br,
                     hs.as<uint32_t>(), hc, st);
    launch_cand_from_keys(k.as<uint64_t>(), v.as<uint32_t>(), n, bt, bw, c, st);
    SPS_HIP(hipStreamSynchronize(st));
}

// Every candidate of the image, sorted by (mdt, mt, mwidth), into `out`
// (grown to the count); returns the count.  Synchronises st.
int64_t gather_sorted(const float* img, int64_t rows, int64_t cols, int64_t pitch, const float* d_thr, float h_thr,
                      int wnd, bool pow2, DevBuf& out, DevBuf& cnt, hipStream_t st) {
    const int64_t wmax = pow2 ? (int64_t)1 << wnd : wnd;
    cnt.ensure(sizeof(unsigned long long));
    int64_t tcap = std::max<int64_t>((int64_t)(out.bytes / sizeof(CandD)), 1024);
    unsigned long long n = 0;
    for (int pass = 0; pass < 2; ++pass) {
        out.ensure(sizeof(CandD) * (size_t)tcap);
        launch_fill_u64(cnt.as<unsigned long long>(), 1, 0ull, st);
        launch_boxcar(img, rows, cols, pitch, d_thr, h_thr, (int)wmax, pow2, out.as<CandD>(), tcap,
                      cnt.as<unsigned long long>(), st);
        SPS_HIP(hipMemcpyAsync(&n, cnt.p, sizeof(n), hipMemcpyDeviceToHost, st));
        SPS_HIP(hipStreamSynchronize(st));
        if ((int64_t)n <= tcap) break;
        tcap = (int64_t)n;
    }
    sort_cands(out.as<CandD>(), (int64_t)n, rows, cols, pow2 ? wnd : wmax, st);
    return (int64_t)n;
}

// Three-key gap clustering (Clusterization.cu:92-166) of n candidates with the
// device bins, delegates sorted by mt into d_delegates / d_group_size;
// returns the group count.  Synchronises st.
int64_t cluster_candidates(const CandD* c, int64_t n, const int32_t* d_bins, int32_t* d_delegates,
                           int64_t* d_group_size, hipStream_t st) {
    if (n == 0) return 0;
    SPS_REQUIRE(n < 0xffffffffLL, "invalid candidate count");
    DevBuf k, kt, v, vt, perm, grp, flags, hs, best, start, sc;
    k.ensure(sizeof(uint64_t) * (size_t)n);
    kt.ensure(sizeof(uint64_t) * (size_t)n);
    v.ensure(sizeof(uint32_t) * (size_t)n);
    vt.ensure(sizeof(uint32_t) * (size_t)n);
    perm.ensure(sizeof(uint32_t) * (size_t)n);
    grp.ensure(sizeof(uint32_t) * (size_t)n);
    flags.ensure(sizeof(uint32_t) * (size_t)n);
    const int64_t hc = radix_hist_entries(n);
    hs.ensure(sizeof(uint32_t) * (size_t)hc);
    sc.ensure(sizeof(uint32_t) * (size_t)scan_scratch_entries(n));
    launch_iota_u32(perm.as<uint32_t>(), n, st);
    for (int level = 0; level < 3; ++level) {
        launch_cluster_keys(c, perm.as<uint32_t>(), level == 0 ? nullptr : grp.as<uint32_t>(), n, level,
                            k.as<uint64_t>(), v.as<uint32_t>(), st);
        radix_sort_pairs(k.as<uint64_t>(), v.as<uint32_t>(), kt.as<uint64_t>(), vt.as<uint32_t>(), n,
                         level == 0 ? 32 : 64, hs.as<uint32_t>(), hc, st);
        SPS_HIP(hipMemcpyAsync(perm.p, v.p, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToDevice, st));
        launch_cluster_flags(k.as<uint64_t>(), n, d_bins, level, flags.as<uint32_t>(), st);
        SPS_HIP(hipMemcpyAsync(grp.p, flags.p, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToDevice, st));
        scan_inclusive_u32(grp.as<uint32_t>(), n, sc.as<uint32_t>(), st);
    }
    uint32_t ng = 0;
    SPS_HIP(hipMemcpyAsync(&ng, grp.as<uint32_t>() + (n - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    SPS_HIP(hipStreamSynchronize(st));
    best.ensure(sizeof(unsigned long long) * (size_t)ng);
    start.ensure(sizeof(uint32_t) * (size_t)ng);
    launch_fill_u64(best.as<unsigned long long>(), ng, 0ull, st);
    launch_cluster_delegates(c, perm.as<uint32_t>(), grp.as<uint32_t>(), flags.as<uint32_t>(), n,
                             best.as<unsigned long long>(), start.as<uint32_t>(), st);
    launch_delegate_keys(c, best.as<unsigned long long>(), start.as<uint32_t>(), ng, n, k.as<uint64_t>(),
                         v.as<uint32_t>(), st);
    radix_sort_pairs(k.as<uint64_t>(), v.as<uint32_t>(), kt.as<uint64_t>(), vt.as<uint32_t>(), ng, 64,
                     hs.as<uint32_t>(), hc, st);
    launch_delegate_out(best.as<unsigned long long>(), start.as<uint32_t>(), v.as<uint32_t>(), ng, n, d_delegates,
                        d_group_size, st);
    SPS_HIP(hipStreamSynchronize(st));
    return (int64_t)ng;
}

}  // namespace

SPS_API sps_status sps_boxcar_candidates(const float* d_img, int64_t rows, int64_t cols, const float* d_thr, int wnd,
                                         int all_widths, sps_cand* d_out, int64_t cap, int64_t* h_count,
                                         void* stream) {
    return guarded([&] {
        SPS_REQUIRE(d_img && d_thr && h_count, "null argument");
        SPS_REQUIRE(rows > 0 && cols > 0 && cap >= 0, "invalid sizes");
        SPS_REQUIRE(cap == 0 || d_out, "null output");
        const bool pow2 = !all_widths;
        SPS_REQUIRE(wnd >= (pow2 ? 0 : 1) && wnd <= (pow2 ? 13 : 8192), "window parameter out of range");
        hipStream_t st = as_stream(stream);
        DevBuf cnt, tmp;
        const int64_t n = gather_sorted(d_img, rows, cols, cols, d_thr, 0.f, wnd, pow2, tmp, cnt, st);
        const int64_t ncopy = std::min<int64_t>(n, cap);
        if (ncopy > 0)
            SPS_HIP(hipMemcpyAsync(d_out, tmp.p, sizeof(CandD) * (size_t)ncopy, hipMemcpyDeviceToDevice, st));
        SPS_HIP(hipStreamSynchronize(st));
        *h_count = n;
        if (n > cap) throw Error(SPS_ERR_CAPACITY, "more candidates than the output capacity");
    });
}

SPS_API sps_status sps_cluster(const sps_cand* d_cands, int64_t n, const int32_t* d_bins, int32_t* d_delegates,
                               int64_t* d_group_size, int64_t* h_ngroups, void* stream) {
    return guarded([&] {
        SPS_REQUIRE(h_ngroups && d_bins, "null argument");
        SPS_REQUIRE(n >= 0 && n < 0xffffffffLL, "invalid candidate count");
        *h_ngroups = 0;
        if (n == 0) return;
        SPS_REQUIRE(d_cands && d_delegates && d_group_size, "null argument");
        *h_ngroups = cluster_candidates(reinterpret_cast<const CandD*>(d_cands), n, d_bins, d_delegates, d_group_size,
                                        as_stream(stream));
    });
}

SPS_API sps_status sps_detect_max(const float* d_arr, const float* d_norm, int64_t rows, int64_t cols, int wnd,
                                  sps_detection* d_out, void* stream) {
    return guarded([&] {
        SPS_REQUIRE(d_arr && d_out && rows > 0 && cols > 0 && wnd >= 1, "invalid argument");
        hipStream_t st = as_stream(stream);
        DevBuf blocks;
        const int64_t nb = detect_max_blocks(rows, cols);
        blocks.ensure(16 * (size_t)nb);
        launch_detect_max(d_arr, d_norm, rows, cols, wnd, blocks.p, nb, d_out, st);
        SPS_HIP(hipStreamSynchronize(st));
    });
}

SPS_API sps_status sps_boxcar_snr(const float* d_img, int64_t rows, int64_t cols, int w, float* d_snr, void* stream) {
    return guarded([&] {
        SPS_REQUIRE(d_img && d_snr && rows > 0 && cols > 0 && w >= 1, "invalid argument");
        launch_boxcar_snr(d_img, rows, cols, w, d_snr, as_stream(stream));
    });
}

SPS_API sps_status sps_digitize(float* d_x, int64_t n, void* stream) {
    return guarded([&] {
        SPS_REQUIRE(n >= 0 && (d_x || n == 0), "invalid argument");
        launch_digitize(d_x, n, as_stream(stream));
    });
}

SPS_API sps_status sps_boxcar_snr_adjoint(const float* d_g, int64_t rows, int64_t cols, int w, float* d_gi,
                                          void* stream) {
    return guarded([&] {
        SPS_REQUIRE(d_g && d_gi && rows > 0 && cols > 0 && w >= 1, "invalid argument");
        launch_boxcar_snr_adjoint(d_g, rows, cols, w, d_gi, as_stream(stream));
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// fused single-beam search
// ---------------------------------------------------------------------------
struct sps_search {
    sps_search_params p{};
    std::unique_ptr<sps_fdmt> f;     // data transform (nsamps)
    int64_t ndt = 0;
    int64_t norm_len = 0;            // samples of the mask transform
    DevBuf rm, rv, stats, mu, sg, keep, zeros, ones_f;
    DevBuf norm_img, norm_a, norm_b; // mask transform output and its states
    DevBuf mask_shift;               // [ndt][nchans] int16 channel shifts (empty: run the transform)
    DevBuf dmt;                      // internal DM-time image when the caller passes none
    bool ready = false;
    // clustering leg (sps_search_cluster): all-widths candidates, delegates
    DevBuf heap, heap_cnt, d_del, d_gsz;
    int64_t heap_n = 0;
    std::vector<int32_t> h_del;
    std::vector<int64_t> h_gsz;
    // stage timing: ev[0] start, ev[1] stats, ev[2..] init + merges, then
    // mask transform, S/N normalisation, boxcar
    bool timing = false;
    std::vector<hipEvent_t> ev;

    ~sps_search() {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
    // stats, mask transform, init, one per pass, separate S/N, boxcar
    int64_t nstages() const { return (int64_t)f->schedule.size() + 5; }
    void ensure_events() {
        const size_t need = (size_t)nstages() + 1;
        while (ev.size() < need) {
            hipEvent_t e;
            SPS_HIP(hipEventCreate(&e));
            ev.push_back(e);
        }
    }

    // device scratch of one search (floats): ping-pong states, mask transform
    // states and image, and the DM-time image when the caller passes none.
    // A batch hands one set, sized for the largest of its beams, to every
    // beam that runs on the same stream.
    struct Scratch {
        float *sa = nullptr, *sb = nullptr, *na = nullptr, *nb = nullptr, *nimg = nullptr, *dmt = nullptr;
    };
    struct ScratchSize {
        int64_t sa = 0, sb = 0, na = 0, nb = 0, nimg = 0, dmt = 0;
    };
    ScratchSize scratch_size() const {
        ScratchSize z;
        z.sa = checked_mul(std::max<int64_t>(f->rows_a, 1), f->pitch, "state");
        z.sb = checked_mul(std::max<int64_t>(f->rows_b, 1), f->pitch, "state");
        if (p.snr_norm) {
            const int64_t Pn = round_up(norm_len, 4);
            z.na = checked_mul(std::max<int64_t>(f->rows_a, 1), Pn, "norm");
            z.nb = checked_mul(std::max<int64_t>(f->rows_b, 1), Pn, "norm");
            z.nimg = checked_mul(ndt, norm_len, "norm");
        }
        z.dmt = checked_mul(ndt, p.nsamps, "dmt");
        return z;
    }

    void setup(hipStream_t st, bool own_scratch = true) {
        if (!ready) setup_small(st);
        if (own_scratch) {
            f->ensure_states();
            if (p.snr_norm) {
                const ScratchSize z = scratch_size();
                norm_img.ensure(sizeof(float) * (size_t)z.nimg);
                norm_a.ensure(sizeof(float) * (size_t)z.na);
                norm_b.ensure(sizeof(float) * (size_t)z.nb);
            }
        }
        f->ensure_trees(st);
    }

    Scratch own_scratch() {
        Scratch s;
        s.sa = f->state_a.as<float>();
        s.sb = f->state_b.as<float>();
        s.na = norm_a.as<float>();
        s.nb = norm_b.as<float>();
        s.nimg = norm_img.as<float>();
        return s;
    }

    // per-search tables and statistics buffers (small)
    void setup_small(hipStream_t st) {
        f->ensure_device(st);
        const int64_t R = p.nchans;
        rm.ensure(sizeof(double) * (size_t)R);
        rv.ensure(sizeof(double) * (size_t)R);
        stats.ensure(sizeof(double) * 8);
        mu.ensure(sizeof(float) * (size_t)R);
        sg.ensure(sizeof(float) * (size_t)R);
        keep.ensure((size_t)R);
        std::vector<float> z((size_t)R, 0.f), o((size_t)R, 1.f);
        upload(zeros, z, st);
        upload(ones_f, o, st);
        if (p.norm_mode == SPS_NORM_NONE) {
            // no cleaning: every channel kept, unit scale, mask of ones
            std::vector<uint8_t> k((size_t)R, 1);
            upload(keep, k, st);
            upload(mu, z, st);
            upload(sg, o, st);
            std::vector<double> s = {0.0, 1.0, 0.0, 0.0, (double)R, 0.0, 0.0, 0.0};
            upload(stats, s, st);
        }
        SPS_HIP(hipStreamSynchronize(st));
        ready = true;
    }

    // Enqueues one beam on st (no host synchronisation, no allocation, no
    // memset: safe under stream capture once setup() has run).
    void enqueue(const void* d_in, float* d_dmt, sps_cand* d_cands, uint64_t* d_count, hipStream_t st,
                 const Scratch& sc, hipEvent_t* ev) {
        const int64_t N = p.nsamps;
        const int it = (int)f->schedule.size();  // passes
        if (ev) SPS_HIP(hipEventRecord(ev[0], st));
        if (p.norm_mode != SPS_NORM_NONE) {
            launch_row_stats(d_in, p.in_dtype == SPS_U8, p.nchans, N, rm.as<double>(), rv.as<double>(), st);
            launch_stats_finalize(rm.as<double>(), rv.as<double>(), p.nchans, p.norm_mode, stats.as<double>(),
                                  mu.as<float>(), sg.as<float>(), keep.as<uint8_t>(), nullptr, st);
        }
        if (ev) SPS_HIP(hipEventRecord(ev[1], st));
        float* out = d_dmt ? d_dmt : sc.dmt;
        SPS_REQUIRE(out, "no DM-time image buffer");
        // mask transform first: its output feeds the S/N epilogue of the last merge
        MergeEpilogue epi{nullptr, 0, nullptr};
        const bool fuse = p.snr_norm && f->eng.niters > 0;
        if (p.snr_norm) {
            const int64_t L = norm_len;
            if (mask_shift.p) {
                launch_mask_counts(mask_shift.as<int16_t>(), keep.as<uint8_t>(), ndt, p.nchans, (int)L, sc.nimg, L,
                                   st);
            } else {
                ChanNorm mask{zeros.as<float>(), ones_f.as<float>(), keep.as<uint8_t>()};
                f->run(nullptr, 0, 1, N, &mask, L, sc.nimg, L, sc.na, sc.nb, round_up(L, 4), st, nullptr, nullptr,
                       true);
            }
            epi = MergeEpilogue{sc.nimg, L, stats.as<double>()};
        }
        if (ev) SPS_HIP(hipEventRecord(ev[2], st));
        ChanNorm cn{mu.as<float>(), sg.as<float>(), keep.as<uint8_t>()};
        const ChanNorm* np = (p.norm_mode != SPS_NORM_NONE) ? &cn : nullptr;
        f->run(d_in, p.in_dtype == SPS_U8, 0, N, np, N, out, N, sc.sa, sc.sb, f->pitch, st, ev ? ev + 3 : nullptr,
               fuse ? &epi : nullptr);
        if (p.snr_norm && !fuse)
            launch_snr_normalise(out, ndt, N, N, sc.nimg, norm_len, stats.as<double>(), st);
        if (ev) SPS_HIP(hipEventRecord(ev[it + 4], st));
        launch_fill_u64(reinterpret_cast<unsigned long long*>(d_count), 1, 0ull, st);
        launch_boxcar(out, ndt, N, N, nullptr, p.threshold, 1 << p.box_q, 1, reinterpret_cast<CandD*>(d_cands),
                      p.cand_capacity, reinterpret_cast<unsigned long long*>(d_count), st);
        if (ev) SPS_HIP(hipEventRecord(ev[it + 5], st));
    }
};

// The mask transform's input is 1 on every sample of a kept channel and 0
// elsewhere.  With the running-mean init its cells are exactly 1 from t = the
// cell's dt on and 0 before (a mean of ones; cells t < dt are zero -- the
// CFdmt init keeps partial means there instead, so it is not counted), and
// each merge adds exact small
// integers, a head counted from t = offset on.  So output row r at t counts
// the kept channels whose accumulated shift along r's tree (init dt plus the
// head offsets on the path) is <= t.  The shifts are tabulated here per
// (output row, channel); a row that would hold a channel twice, or shifts
// beyond int16, leave the table empty and the transform runs instead.
static void build_mask_shifts(sps_search& s) {
    const EnginePlan& e = s.f->eng;
    const int64_t L = s.norm_len;
    const int64_t R = e.out_rows, C = e.nchans;
    // L + 256 LDS counters within 64 KiB
    if (e.init_mode != kInitMean || L <= 0 || L > 16000 || R <= 0 || C <= 0 || R * C > ((int64_t)1 << 31)) return;
    // state-0 row -> channel and init dt
    std::vector<int32_t> chan0((size_t)e.rows[0], -1), dt0((size_t)e.rows[0], 0);
    for (int64_t c = 0; c < C; ++c)
        for (int64_t k = e.grid_off[(size_t)c]; k < e.grid_off[(size_t)c + 1]; ++k) {
            const int64_t r0 = e.chan_row_off[(size_t)c] + (k - e.grid_off[(size_t)c]);
            chan0[(size_t)r0] = (int32_t)c;
            dt0[(size_t)r0] = e.grid[(size_t)k];
        }
    // per iteration: output row -> its record
    std::vector<std::vector<int64_t>> rec((size_t)e.niters + 1);
    for (int i = 1; i <= e.niters; ++i) {
        rec[(size_t)i].assign((size_t)e.rows[(size_t)i], -1);
        const auto& it = e.iter[(size_t)i];
        for (size_t j = 0; j < it.size(); ++j) rec[(size_t)i][(size_t)it[j].out_row] = (int64_t)j;
    }
    std::vector<int16_t> tab((size_t)(R * C), (int16_t)32767);
    struct Node {
        int i;
        int64_t row, base;
    };
    std::vector<Node> stack;
    for (int64_t r = 0; r < R; ++r) {
        int16_t* trow = tab.data() + (size_t)(r * C);
        stack.assign(1, Node{e.niters, r, 0});
        while (!stack.empty()) {
            const Node n = stack.back();
            stack.pop_back();
            if (n.i == 0) {
                const int32_t c = chan0[(size_t)n.row];
                const int64_t sh = n.base + dt0[(size_t)n.row];
                if (c < 0 || trow[c] != 32767 || sh < 0 || sh >= 32767) return;  // unexpected: keep the transform
                trow[c] = (int16_t)sh;
                continue;
            }
            const int64_t j = rec[(size_t)n.i][(size_t)n.row];
            if (j < 0) return;
            const RowDesc& d = e.iter[(size_t)n.i][(size_t)j];
            stack.push_back(Node{n.i - 1, d.tail_row, n.base});
            if (d.head_row >= 0) stack.push_back(Node{n.i - 1, d.head_row, n.base + d.offset});
        }
    }
    s.mask_shift.ensure(sizeof(int16_t) * tab.size());
    SPS_HIP(hipMemcpy(s.mask_shift.p, tab.data(), sizeof(int16_t) * tab.size(), hipMemcpyHostToDevice));
}

// cfdmt_imaxdt > 0: the data transform is the legacy CFdmt plan
// (CFdmtGpu(f_min, f_max, nchans, nsamps, imaxDt)) instead of the FDMT plan;
// dt_max and tsamp are then unused
static std::unique_ptr<sps_search> make_search(const sps_search_params* p, int cfdmt_imaxdt = 0) {
    SPS_REQUIRE(p->in_dtype == SPS_F32 || p->in_dtype == SPS_U8, "unknown dtype");
    SPS_REQUIRE(p->norm_mode >= SPS_NORM_NONE && p->norm_mode <= SPS_NORM_PER_CHANNEL, "unknown norm mode");
    SPS_REQUIRE(p->box_q >= 0 && p->box_q <= 13, "box_q out of range");
    SPS_REQUIRE(p->cand_capacity >= 0, "negative capacity");
    std::unique_ptr<sps_search> s(new sps_search());
    s->p = *p;
    s->f.reset(new sps_fdmt());
    if (cfdmt_imaxdt > 0) {
        SPS_REQUIRE(p->nchans <= INT32_MAX && p->nsamps <= INT32_MAX, "CFdmt sizes must fit in int");
        s->f->cplan.reset(new CFdmtPlan(p->f_min, p->f_max, (int)p->nchans, (int)p->nsamps, cfdmt_imaxdt));
        s->f->eng = s->f->cplan->lower();
    } else {
        s->f->plan.reset(new FdmtPlan(p->f_min, p->f_max, p->nchans, p->nsamps, p->tsamp, p->dt_max, 1, 0));
        s->f->eng = s->f->plan->lower(kInitMean);
    }
    s->f->init_layout();
    s->ndt = s->f->eng.out_rows;
    // the mask transform is causal: samples t < L depend on inputs before
    // L only, and beyond the longest channel track every kept channel
    // contributes exactly 1, so the divisor there is the kept count
    const EnginePlan& e = s->f->eng;
    std::vector<int64_t> ext((size_t)e.rows[0]);
    for (int64_t c = 0; c < e.nchans; ++c)
        for (int64_t k = e.grid_off[(size_t)c]; k < e.grid_off[(size_t)c + 1]; ++k)
            ext[(size_t)(e.chan_row_off[(size_t)c] + (k - e.grid_off[(size_t)c]))] = e.grid[(size_t)k];
    for (int i = 1; i <= e.niters; ++i) {
        std::vector<int64_t> nx((size_t)e.rows[(size_t)i], 0);
        for (const auto& r : e.iter[(size_t)i]) {
            int64_t v = ext[(size_t)r.tail_row];
            if (r.head_row >= 0) v = std::max(v, ext[(size_t)r.head_row] + r.offset);
            nx[(size_t)r.out_row] = v;
        }
        ext.swap(nx);
    }
    int64_t mx = 0;
    for (auto v : ext) mx = std::max(mx, v);
    s->norm_len = std::min<int64_t>(p->nsamps, mx + 1);
    build_mask_shifts(*s);
    return s;
}

extern "C" {

SPS_API sps_status sps_search_create(const sps_search_params* p, sps_search** out) {
    return guarded([&] {
        SPS_REQUIRE(p && out, "null argument");
        *out = nullptr;
        *out = make_search(p).release();
    });
}

SPS_API sps_status sps_search_create_cfdmt(const sps_search_params* p, int imaxdt, sps_search** out) {
    return guarded([&] {
        SPS_REQUIRE(p && out, "null argument");
        SPS_REQUIRE(imaxdt >= 1, "imaxdt must be positive");
        *out = nullptr;
        *out = make_search(p, imaxdt).release();
    });
}

SPS_API sps_status sps_search_destroy(sps_search* s) {
    return guarded([&] { delete s; });
}

SPS_API sps_status sps_search_get_ndt(const sps_search* s, int64_t* n) {
    return guarded([&] {
        SPS_REQUIRE(s && n, "null argument");
        *n = s->ndt;
    });
}

SPS_API sps_status sps_search_run(sps_search* s, const void* d_in, float* d_dmt, sps_cand* d_cands,
                                  uint64_t* d_count, void* stream) {
    return guarded([&] {
        SPS_REQUIRE(s && d_in && d_count, "null argument");
        SPS_REQUIRE(s->p.cand_capacity == 0 || d_cands, "null candidate buffer");
        hipStream_t st = as_stream(stream);
        s->setup(st);
        sps_search::Scratch sc = s->own_scratch();
        if (!d_dmt) {
            s->dmt.ensure(sizeof(float) * (size_t)s->scratch_size().dmt);
            sc.dmt = s->dmt.as<float>();
        }
        hipEvent_t* ev = nullptr;
        if (s->timing) {
            s->ensure_events();
            ev = s->ev.data();
        }
        s->enqueue(d_in, d_dmt, d_cands, d_count, st, sc, ev);
    });
}

SPS_API sps_status sps_fdmt_set_schedule(sps_fdmt* h, const int32_t* h_depths, int64_t n) {
    return guarded([&] {
        SPS_REQUIRE(h && (n == 0 || h_depths), "null argument");
        h->set_schedule(std::vector<int>(h_depths, h_depths + n));
    });
}

SPS_API sps_status sps_fdmt_get_schedule(const sps_fdmt* h, int32_t* h_depths, int64_t cap, int64_t* h_n) {
    return guarded([&] {
        SPS_REQUIRE(h && h_n, "null argument");
        *h_n = (int64_t)h->schedule.size();
        for (int64_t i = 0; i < *h_n && i < cap && h_depths; ++i) h_depths[i] = h->schedule[(size_t)i];
    });
}

SPS_API sps_status sps_fdmt_set_staged(sps_fdmt* h, int staged) {
    return guarded([&] {
        SPS_REQUIRE(h, "null handle");
        h->set_staged(staged);
    });
}

SPS_API sps_status sps_search_set_staged(sps_search* s, int staged) {
    return guarded([&] {
        SPS_REQUIRE(s, "null handle");
        s->f->set_staged(staged);
    });
}

SPS_API sps_status sps_fdmt_set_group_shape(sps_fdmt* h, int lds_floats, int max_rows, int tiles_per_block) {
    return guarded([&] {
        SPS_REQUIRE(h, "null handle");
        h->set_group_shape(lds_floats, max_rows, tiles_per_block);
    });
}

SPS_API sps_status sps_search_set_group_shape(sps_search* s, int lds_floats, int max_rows, int tiles_per_block) {
    return guarded([&] {
        SPS_REQUIRE(s, "null handle");
        s->f->set_group_shape(lds_floats, max_rows, tiles_per_block);
    });
}

SPS_API sps_status sps_search_set_schedule(sps_search* s, const int32_t* h_depths, int64_t n) {
    return guarded([&] {
        SPS_REQUIRE(s && (n == 0 || h_depths), "null argument");
        s->f->set_schedule(std::vector<int>(h_depths, h_depths + n));
    });
}

SPS_API sps_status sps_search_get_schedule(const sps_search* s, int32_t* h_depths, int64_t cap, int64_t* h_n) {
    return guarded([&] {
        SPS_REQUIRE(s && h_n, "null argument");
        *h_n = (int64_t)s->f->schedule.size();
        for (int64_t i = 0; i < *h_n && i < cap && h_depths; ++i) h_depths[i] = s->f->schedule[(size_t)i];
    });
}

SPS_API sps_status sps_search_enable_timing(sps_search* s, int enable) {
    return guarded([&] {
        SPS_REQUIRE(s, "null handle");
        s->timing = enable != 0;
        if (s->timing) s->ensure_events();
    });
}

SPS_API sps_status sps_search_get_timing(sps_search* s, double* h_ms, int64_t cap, int64_t* h_n) {
    return guarded([&] {
        SPS_REQUIRE(s && h_n, "null argument");
        SPS_REQUIRE(s->timing, "timing is not enabled");
        const int64_t n = s->nstages();
        *h_n = n;
        SPS_HIP(hipEventSynchronize(s->ev[(size_t)n]));
        for (int64_t i = 0; i < n && i < cap; ++i) {
            float ms = 0.f;
            SPS_HIP(hipEventElapsedTime(&ms, s->ev[(size_t)i], s->ev[(size_t)i + 1]));
            h_ms[i] = ms;
        }
    });
}

SPS_API sps_status sps_search_get_stage_bytes(const sps_search* s, double* h_b, int64_t cap, int64_t* h_n) {
    return guarded([&] {
        SPS_REQUIRE(s && h_n, "null argument");
        const sps_fdmt& f = *s->f;
        const EnginePlan& e = f.eng;
        const double N = (double)s->p.nsamps, N4 = 4.0 * N;
        const double in_elem = s->p.in_dtype == SPS_U8 ? 1.0 : 4.0;
        const double in_bytes = in_elem * N * (double)e.nchans;
        std::vector<double> b;
        b.push_back(s->p.norm_mode != SPS_NORM_NONE ? in_bytes : 0.0);  // statistics: one input read
        // mask transform: its output image (ndt rows over its extent), written
        // once and read once by the final pass; its states stay in the caches
        b.push_back(s->p.snr_norm ? 2.0 * 4.0 * (double)s->norm_len * (double)s->ndt : 0.0);
        // init: input read + live state-0 rows; with the init fused into the
        // first pass, state 0 never reaches HBM and that pass reads the input
        b.push_back(f.ran_fused0 ? 0.0 : in_bytes + N4 * (double)f.live_rows(0));
        int i0 = 1;
        for (int L : f.schedule) {  // a pass reads the live rows of state i0-1 and writes those of state i0+L-1
            const double rd = (i0 == 1 && f.ran_fused0) ? in_bytes : N4 * (double)f.live_rows(i0 - 1);
            b.push_back(rd + N4 * (double)f.live_rows(i0 + L - 1));
            i0 += L;
        }
        b.push_back(0.0);                                                // S/N normalisation (fused into the last pass)
        b.push_back(N4 * (double)s->ndt);                                // boxcar: one read of the image
        *h_n = (int64_t)b.size();
        for (int64_t i = 0; i < (int64_t)b.size() && i < cap; ++i) h_b[i] = b[(size_t)i];
    });
}

SPS_API sps_status sps_search_sort_candidates(sps_search* s, sps_cand* d_cands, const uint64_t* d_count,
                                              int64_t* h_count, void* stream) {
    return guarded([&] {
        SPS_REQUIRE(s && d_count && h_count, "null argument");
        hipStream_t st = as_stream(stream);
        unsigned long long n = 0;
        SPS_HIP(hipMemcpyAsync(&n, d_count, sizeof(n), hipMemcpyDeviceToHost, st));
        SPS_HIP(hipStreamSynchronize(st));
        const int64_t m = std::min<int64_t>((int64_t)n, s->p.cand_capacity);
        sort_cands(reinterpret_cast<CandD*>(d_cands), m, s->ndt, s->p.nsamps, s->p.box_q, st);
        *h_count = (int64_t)n;
        // the kept prefix is whichever candidates won the counter first: the
        // caller must re-gather into a buffer of h_count entries
        if ((int64_t)n > s->p.cand_capacity)
            throw Error(SPS_ERR_CAPACITY, "more candidates than cand_capacity: re-gather with sps_search_gather");
    });
}

SPS_API sps_status sps_search_gather(sps_search* s, const float* d_dmt, sps_cand* d_cands, int64_t cap,
                                     uint64_t* d_count, int64_t* h_count, void* stream) {
    return guarded([&] {
        SPS_REQUIRE(s && d_count && h_count && cap >= 0 && (cap == 0 || d_cands), "null argument");
        const float* img = d_dmt ? d_dmt : s->dmt.as<float>();
        SPS_REQUIRE(img, "no DM-time image: run the search first (or pass the image it wrote)");
        hipStream_t st = as_stream(stream);
        const int64_t N = s->p.nsamps;
        launch_fill_u64(reinterpret_cast<unsigned long long*>(d_count), 1, 0ull, st);
        launch_boxcar(img, s->ndt, N, N, nullptr, s->p.threshold, 1 << s->p.box_q, 1,
                      reinterpret_cast<CandD*>(d_cands), cap, reinterpret_cast<unsigned long long*>(d_count), st);
        unsigned long long n = 0;
        SPS_HIP(hipMemcpyAsync(&n, d_count, sizeof(n), hipMemcpyDeviceToHost, st));
        SPS_HIP(hipStreamSynchronize(st));
        sort_cands(reinterpret_cast<CandD*>(d_cands), std::min<int64_t>((int64_t)n, cap), s->ndt, N, s->p.box_q, st);
        *h_count = (int64_t)n;
        if ((int64_t)n > cap) throw Error(SPS_ERR_CAPACITY, "more candidates than the gather's capacity");
    });
}

SPS_API sps_status sps_search_cluster(sps_search* s, const float* d_dmt, int wnd, const int32_t* d_bins,
                                      int64_t* h_ncand, int64_t* h_ngroups, void* stream) {
    return guarded([&] {
        SPS_REQUIRE(s && d_bins && h_ncand && h_ngroups, "null argument");
        SPS_REQUIRE(wnd >= 1 && wnd <= 8192, "window must be 1..8192");
        const float* img = d_dmt ? d_dmt : s->dmt.as<float>();
        SPS_REQUIRE(img, "no DM-time image: run the search first (or pass the image it wrote)");
        hipStream_t st = as_stream(stream);
        s->heap_n = 0;
        s->h_del.clear();
        s->h_gsz.clear();
        *h_ncand = 0;
        *h_ngroups = 0;
        const int64_t N = s->p.nsamps;
        const int64_t n = gather_sorted(img, s->ndt, N, N, nullptr, s->p.threshold, wnd, false, s->heap, s->heap_cnt, st);
        s->heap_n = n;
        *h_ncand = n;
        if (n == 0) return;
        s->d_del.ensure(sizeof(int32_t) * (size_t)n);
        s->d_gsz.ensure(sizeof(int64_t) * (size_t)n);
        const int64_t ng = cluster_candidates(s->heap.as<CandD>(), n, d_bins, s->d_del.as<int32_t>(),
                                              s->d_gsz.as<int64_t>(), st);
        s->h_del.resize((size_t)ng);
        s->h_gsz.resize((size_t)ng);
        if (ng > 0) {
            SPS_HIP(hipMemcpyAsync(s->h_del.data(), s->d_del.p, sizeof(int32_t) * (size_t)ng, hipMemcpyDeviceToHost, st));
            SPS_HIP(hipMemcpyAsync(s->h_gsz.data(), s->d_gsz.p, sizeof(int64_t) * (size_t)ng, hipMemcpyDeviceToHost, st));
        }
        SPS_HIP(hipStreamSynchronize(st));
        *h_ngroups = ng;
    });
}

SPS_API sps_status sps_search_get_delegates(const sps_search* s, sps_cand* h_delegates, int64_t* h_group_sizes,
                                            int64_t cap, int64_t* h_n) {
    return guarded([&] {
        SPS_REQUIRE(s && h_n && cap >= 0, "null argument");
        const int64_t ng = (int64_t)s->h_del.size();
        *h_n = ng;
        const int64_t m = std::min(ng, cap);
        for (int64_t i = 0; i < m; ++i) {
            if (h_delegates)
                SPS_HIP(hipMemcpy(&h_delegates[i], s->heap.as<CandD>() + s->h_del[(size_t)i], sizeof(CandD),
                                  hipMemcpyDeviceToHost));
            if (h_group_sizes) h_group_sizes[i] = s->h_gsz[(size_t)i];
        }
    });
}

SPS_API sps_status sps_search_write_log(const sps_search* s, const char* filename, float delt_bin,
                                        float t_chunk_begin) {
    if (!s || !filename) {
        g_last_error = "null argument";
        return SPS_ERR_INVALID_ARGUMENT;
    }
    return sps_write_log(reinterpret_cast<const sps_cand*>(s->heap.p), s->heap_n, s->h_del.data(), s->h_gsz.data(),
                         (int64_t)s->h_del.size(), filename, delt_bin, t_chunk_begin);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// batched multi-beam search
//
// Every beam is a full search (its own plan, tables and statistics).  Beams
// are spread over a few worker streams by longest-first greedy assignment on
// their byte cost; the beams of one worker run one after the other and share
// one set of scratch buffers sized for the largest of them, so the HBM held
// is the inputs, the candidate buffers and nworkers x the largest beam's
// states.  The whole fork/join launch sequence is captured once into a HIP
// graph and replayed; it holds kernels and event waits only.
// ---------------------------------------------------------------------------
struct sps_batch {
    struct Beam {
        sps_batch_beam io{};
        std::unique_ptr<sps_search> s;
        int worker = 0;
        double cost = 0.0;
    };
    struct Worker {
        hipStream_t st = nullptr;
        hipEvent_t join = nullptr;
        DevBuf sa, sb, na, nb, nimg, dmt;
        sps_search::Scratch sc;
        std::vector<int64_t> beams;
    };
    sps_batch_params p{};
    std::vector<Beam> beams;
    std::vector<Worker> workers;
    hipEvent_t fork = nullptr;
    hipStream_t cap = nullptr;  // capture origin
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool prepared = false;

    ~sps_batch() {
        drop_graph();
        for (auto& w : workers) {
            if (w.join) (void)hipEventDestroy(w.join);
            if (w.st) (void)hipStreamDestroy(w.st);
        }
        if (fork) (void)hipEventDestroy(fork);
        if (cap) (void)hipStreamDestroy(cap);
    }

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }

    void assign() {
        std::vector<int64_t> order(beams.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = (int64_t)i;
        std::stable_sort(order.begin(), order.end(),
                         [&](int64_t a, int64_t b) { return beams[(size_t)a].cost > beams[(size_t)b].cost; });
        std::vector<double> load(workers.size(), 0.0);
        for (auto& w : workers) w.beams.clear();
        for (int64_t b : order) {
            const size_t w = (size_t)(std::min_element(load.begin(), load.end()) - load.begin());
            load[w] += beams[(size_t)b].cost;
            beams[(size_t)b].worker = (int)w;
        }
        // each worker runs its beams in index order
        for (size_t b = 0; b < beams.size(); ++b) workers[(size_t)beams[b].worker].beams.push_back((int64_t)b);
    }

    void prepare(hipStream_t st) {
        if (prepared) return;
        for (auto& w : workers) {
            sps_search::ScratchSize m;
            bool need_dmt = false;
            for (int64_t b : w.beams) {
                const auto z = beams[(size_t)b].s->scratch_size();
                m.sa = std::max(m.sa, z.sa);
                m.sb = std::max(m.sb, z.sb);
                m.na = std::max(m.na, z.na);
                m.nb = std::max(m.nb, z.nb);
                m.nimg = std::max(m.nimg, z.nimg);
                if (!beams[(size_t)b].io.d_dmt) {
                    m.dmt = std::max(m.dmt, z.dmt);
                    need_dmt = true;
                }
            }
            if (w.beams.empty()) continue;
            w.sa.ensure(sizeof(float) * (size_t)m.sa);
            w.sb.ensure(sizeof(float) * (size_t)m.sb);
            if (p.snr_norm) {
                w.na.ensure(sizeof(float) * (size_t)m.na);
                w.nb.ensure(sizeof(float) * (size_t)m.nb);
                w.nimg.ensure(sizeof(float) * (size_t)m.nimg);
            }
            if (need_dmt) w.dmt.ensure(sizeof(float) * (size_t)m.dmt);
            w.sc.sa = w.sa.as<float>();
            w.sc.sb = w.sb.as<float>();
            w.sc.na = w.na.as<float>();
            w.sc.nb = w.nb.as<float>();
            w.sc.nimg = w.nimg.as<float>();
            w.sc.dmt = w.dmt.as<float>();
        }
        for (auto& b : beams) b.s->setup(st, false);
        SPS_HIP(hipStreamSynchronize(st));
        prepared = true;
    }

    // fork on `origin`, every worker runs its beams, join back on `origin`
    void enqueue(hipStream_t origin) {
        SPS_HIP(hipEventRecord(fork, origin));
        for (auto& w : workers) {
            if (w.beams.empty()) continue;
            SPS_HIP(hipStreamWaitEvent(w.st, fork, 0));
            for (int64_t bi : w.beams) {
                Beam& b = beams[(size_t)bi];
                b.s->enqueue(b.io.d_in, b.io.d_dmt, b.io.d_cands, b.io.d_count, w.st, w.sc, nullptr);
            }
            SPS_HIP(hipEventRecord(w.join, w.st));
            SPS_HIP(hipStreamWaitEvent(origin, w.join, 0));
        }
    }

    void capture() {
        drop_graph();
        SPS_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
        try {
            enqueue(cap);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(cap, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        SPS_HIP(hipStreamEndCapture(cap, &graph));
        SPS_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }
};

extern "C" {

SPS_API sps_status sps_batch_create(const sps_batch_params* p, const sps_batch_beam* beams, int64_t nbeams,
                                    sps_batch** out) {
    return guarded([&] {
        SPS_REQUIRE(p && out && (nbeams == 0 || beams), "null argument");
        *out = nullptr;
        SPS_REQUIRE(nbeams >= 1, "a batch needs at least one beam");
        SPS_REQUIRE(p->nworkers >= 1 && p->nworkers <= 8, "nworkers must be 1..8");
        std::unique_ptr<sps_batch> b(new sps_batch());
        b->p = *p;
        b->beams.resize((size_t)nbeams);
        for (int64_t i = 0; i < nbeams; ++i) {
            const sps_batch_beam& io = beams[i];
            SPS_REQUIRE(io.d_in && io.d_count, "null beam buffer");
            SPS_REQUIRE(io.cand_capacity >= 0 && (io.cand_capacity == 0 || io.d_cands), "bad candidate buffer");
            sps_search_params sp{};
            sp.f_min = io.f_min;
            sp.f_max = io.f_max;
            sp.tsamp = io.tsamp;
            sp.nchans = io.nchans;
            sp.nsamps = io.nsamps;
            sp.dt_max = io.dt_max;
            sp.in_dtype = p->in_dtype;
            sp.norm_mode = p->norm_mode;
            sp.snr_norm = p->snr_norm;
            sp.box_q = p->box_q;
            sp.threshold = p->threshold;
            sp.cand_capacity = io.cand_capacity;
            sps_batch::Beam& bm = b->beams[(size_t)i];
            bm.io = io;
            bm.s = make_search(&sp);
            bm.s->f->pc_fast = p->nworkers <= 1;
            // byte cost of the beam: input read twice, every state written
            // once and read once, the image read by the boxcar
            const EnginePlan& e = bm.s->f->eng;
            double rows = (double)e.nchans * 2.0 + (double)e.rows[0];
            for (int k = 1; k <= e.niters; ++k) rows += (double)(e.rows[(size_t)k - 1] + e.rows[(size_t)k]);
            bm.cost = rows * (double)e.nsamps;
        }
        b->workers = std::vector<sps_batch::Worker>((size_t)std::min<int64_t>(p->nworkers, nbeams));
        for (auto& w : b->workers) {
            SPS_HIP(hipStreamCreateWithFlags(&w.st, hipStreamNonBlocking));
            SPS_HIP(hipEventCreateWithFlags(&w.join, hipEventDisableTiming));
        }
        SPS_HIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
        SPS_HIP(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
        b->assign();
        *out = b.release();
    });
}

SPS_API sps_status sps_batch_destroy(sps_batch* b) {
    return guarded([&] { delete b; });
}

SPS_API sps_status sps_batch_get_beam_info(const sps_batch* b, int64_t beam, int64_t* h_ndt, int32_t* h_worker,
                                           int64_t* h_niters) {
    return guarded([&] {
        SPS_REQUIRE(b, "null handle");
        SPS_REQUIRE(beam >= 0 && beam < (int64_t)b->beams.size(), "beam index out of range");
        const auto& bm = b->beams[(size_t)beam];
        if (h_ndt) *h_ndt = bm.s->ndt;
        if (h_worker) *h_worker = bm.worker;
        if (h_niters) *h_niters = bm.s->f->eng.niters;
    });
}

SPS_API sps_status sps_batch_set_schedule(sps_batch* b, int64_t beam, const int32_t* h_depths, int64_t n) {
    return guarded([&] {
        SPS_REQUIRE(b && (n == 0 || h_depths), "null argument");
        SPS_REQUIRE(beam >= 0 && beam < (int64_t)b->beams.size(), "beam index out of range");
        b->beams[(size_t)beam].s->f->set_schedule(std::vector<int>(h_depths, h_depths + n));
        b->drop_graph();
        b->prepared = false;  // the new pass tables are built before the next capture
    });
}

SPS_API sps_status sps_batch_set_group_shape(sps_batch* b, int64_t beam, int lds_floats, int max_rows, int tiles_per_block) {
    return guarded([&] {
        SPS_REQUIRE(b, "null handle");
        SPS_REQUIRE(beam >= 0 && beam < (int64_t)b->beams.size(), "beam index out of range");
        b->beams[(size_t)beam].s->f->set_group_shape(lds_floats, max_rows, tiles_per_block);
        b->drop_graph();
        b->prepared = false;
    });
}

SPS_API sps_status sps_batch_set_staged(sps_batch* b, int64_t beam, int staged) {
    return guarded([&] {
        SPS_REQUIRE(b, "null handle");
        SPS_REQUIRE(beam >= 0 && beam < (int64_t)b->beams.size(), "beam index out of range");
        b->beams[(size_t)beam].s->f->set_staged(staged);
        b->drop_graph();
        b->prepared = false;
    });
}

SPS_API sps_status sps_batch_get_schedule(const sps_batch* b, int64_t beam, int32_t* h_depths, int64_t cap,
                                          int64_t* h_n) {
    return guarded([&] {
        SPS_REQUIRE(b && h_n, "null argument");
        SPS_REQUIRE(beam >= 0 && beam < (int64_t)b->beams.size(), "beam index out of range");
        const auto& s = b->beams[(size_t)beam].s->f->schedule;
        *h_n = (int64_t)s.size();
        for (int64_t i = 0; i < *h_n && i < cap && h_depths; ++i) h_depths[i] = s[(size_t)i];
    });
}

SPS_API sps_status sps_batch_run(sps_batch* b, void* stream) {
    return guarded([&] {
        SPS_REQUIRE(b, "null handle");
        hipStream_t st = as_stream(stream);
        b->prepare(st);
        b->enqueue(st);
    });
}

SPS_API sps_status sps_batch_launch(sps_batch* b, void* stream) {
    return guarded([&] {
        SPS_REQUIRE(b, "null handle");
        hipStream_t st = as_stream(stream);
        b->prepare(st);
        if (!b->exec) b->capture();
        SPS_HIP(hipGraphLaunch(b->exec, st));
    });
}

SPS_API sps_status sps_batch_sort_candidates(sps_batch* b, int64_t beam, int64_t* h_count, void* stream) {
    return guarded([&] {
        SPS_REQUIRE(b && h_count, "null argument");
        SPS_REQUIRE(beam >= 0 && beam < (int64_t)b->beams.size(), "beam index out of range");
        auto& bm = b->beams[(size_t)beam];
        hipStream_t st = as_stream(stream);
        unsigned long long n = 0;
        SPS_HIP(hipMemcpyAsync(&n, bm.io.d_count, sizeof(n), hipMemcpyDeviceToHost, st));
        SPS_HIP(hipStreamSynchronize(st));
        const int64_t m = std::min<int64_t>((int64_t)n, bm.io.cand_capacity);
        sort_cands(reinterpret_cast<CandD*>(bm.io.d_cands), m, bm.s->ndt, bm.s->p.nsamps, bm.s->p.box_q, st);
        *h_count = (int64_t)n;
        if ((int64_t)n > bm.io.cand_capacity)
            throw Error(SPS_ERR_CAPACITY, "more candidates than the beam's cand_capacity");
    });
}

}  // extern "C"
