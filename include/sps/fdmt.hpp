// sps/fdmt.hpp -- drop-in C++ FDMT classes over the libsps C ABI.
//
// Mirrors FDMT_PRAVIR_RESTRICT/fdmt_base.hpp:8-85 (FDMT, FDMTPlan and its
// types), fdmt_gpu.cuh:68-89 (FDMTGPU) and FDMT/FdmtB.h:6-59 +
// FDMT/FdmtGpu.cuh:5-57 (CFdmtB, CFdmtGpu).  Pointers passed to execute /
// process_image are HIP device pointers; work is issued on the default stream
// like the reference.  Errors throw the reference's exception types:
// std::invalid_argument for size/argument errors, std::runtime_error otherwise.
#pragma once

#include <array>
#include <cstddef>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../sps.h"

namespace sps_detail {
inline void check(sps_status s) {
    if (s == SPS_OK) return;
    if (s == SPS_ERR_INVALID_ARGUMENT) throw std::invalid_argument(sps_last_error());
    throw std::runtime_error(sps_last_error());
}
}  // namespace sps_detail

using SizeType = size_t;
using DtGridType = std::vector<SizeType>;
using StShapeType = std::array<SizeType, 5>;      // nchans, ndt_min, ndt_max, ncoords, nsamps
using FDMTCoordType = std::pair<SizeType, SizeType>;  // i_sub, i_dt

struct FDMTCoordMapping {
    FDMTCoordType head;
    FDMTCoordType tail;
    SizeType offset;
};

struct FDMTPlan {
    std::vector<float> df_top;
    std::vector<float> df_bot;
    std::vector<StShapeType> state_shape;
    std::vector<std::vector<FDMTCoordType>> coordinates;
    std::vector<std::vector<FDMTCoordType>> coordinates_to_copy;
    std::vector<std::vector<FDMTCoordMapping>> mappings;
    std::vector<std::vector<FDMTCoordMapping>> mappings_to_copy;
    std::vector<std::vector<SizeType>> state_sub_idx;  // element offsets of each sub-band
    std::vector<std::vector<DtGridType>> dt_grid;
    std::vector<DtGridType> dt_grid_sub_top;

    size_t calculate_memory_usage() const {
        size_t m = (df_top.size() + df_bot.size()) * sizeof(float) + state_shape.size() * sizeof(StShapeType);
        for (const auto& c : coordinates) m += c.size() * sizeof(FDMTCoordType);
        for (const auto& c : mappings) m += c.size() * sizeof(FDMTCoordMapping);
        for (const auto& c : state_sub_idx) m += c.size() * sizeof(SizeType);
        for (const auto& it : dt_grid)
            for (const auto& g : it) m += g.size() * sizeof(SizeType);
        for (const auto& g : dt_grid_sub_top) m += g.size() * sizeof(SizeType);
        return m;
    }
};

// FDMT (fdmt_base.hpp:37-85): the plan lives in libsps; this object owns the handle.
class FDMT {
public:
    FDMT(float f_min, float f_max, SizeType nchans, SizeType nsamps, float tsamp, SizeType dt_max,
         SizeType dt_step = 1, SizeType dt_min = 0, int init_mode = SPS_INIT_MEAN)
        : m_f_min(f_min), m_f_max(f_max), m_nchans(nchans), m_nsamps(nsamps), m_tsamp(tsamp) {
        sps_fdmt* h = nullptr;
        sps_detail::check(sps_fdmt_create(f_min, f_max, (int64_t)nchans, (int64_t)nsamps, tsamp, (int64_t)dt_max,
                                          (int64_t)dt_step, (int64_t)dt_min, init_mode, &h));
        m_h.reset(h, [](sps_fdmt* p) { sps_fdmt_destroy(p); });
        load_plan();
    }
    FDMT(const FDMT&) = delete;
    FDMT& operator=(const FDMT&) = delete;
    FDMT(FDMT&&) = delete;
    FDMT& operator=(FDMT&&) = delete;
    virtual ~FDMT() = default;

    float get_df() const {
        float df, c;
        sps_detail::check(sps_fdmt_get_df(m_h.get(), &df, &c));
        return df;
    }
    float get_correction() const {
        float df, c;
        sps_detail::check(sps_fdmt_get_df(m_h.get(), &df, &c));
        return c;
    }
    int get_m_nsamps() const { return (int)m_nsamps; }
    SizeType get_niters() const { return m_plan.state_shape.size() - 1; }
    const FDMTPlan& get_plan() const { return m_plan; }
    const DtGridType& get_dt_grid_final() const { return m_plan.dt_grid[get_niters()][0]; }
    std::vector<float> get_dm_grid_final() const {
        int64_t n = 0;
        sps_detail::check(sps_fdmt_get_dm_grid_final(m_h.get(), nullptr, 0, &n));
        std::vector<float> v((size_t)n);
        sps_detail::check(sps_fdmt_get_dm_grid_final(m_h.get(), v.data(), n, &n));
        return v;
    }
    virtual void execute(const float* waterfall, size_t waterfall_size, float* dmt, size_t dmt_size) = 0;
    virtual void initialise(const float* waterfall, float* state) = 0;

    sps_fdmt* handle() const { return m_h.get(); }

protected:
    void check_inputs(size_t waterfall_size, size_t dmt_size) const {
        if (waterfall_size != m_nchans * m_nsamps) throw std::invalid_argument("Invalid size of waterfall");
        const auto& s = m_plan.state_shape[get_niters()];
        if (dmt_size != s[3] * s[4]) throw std::invalid_argument("Invalid size of dmt");
    }

private:
    float m_f_min, m_f_max;
    SizeType m_nchans, m_nsamps;
    float m_tsamp;
    std::shared_ptr<sps_fdmt> m_h;
    FDMTPlan m_plan;

    void load_plan() {
        int64_t niters = 0;
        sps_detail::check(sps_fdmt_get_niters(m_h.get(), &niters));
        auto& p = m_plan;
        DtGridType top;
        for (int64_t i = 0; i <= niters; ++i) {
            int64_t s5[5];
            sps_detail::check(sps_fdmt_get_state_shape(m_h.get(), i, s5));
            p.state_shape.push_back({(SizeType)s5[0], (SizeType)s5[1], (SizeType)s5[2], (SizeType)s5[3], (SizeType)s5[4]});
            int64_t nsub = 0, nmap = 0, ncopy = 0;
            float dtop = 0, dbot = 0;
            sps_detail::check(sps_fdmt_get_plan_iter(m_h.get(), i, &nsub, nullptr, nullptr, &nmap, nullptr, &ncopy,
                                                     nullptr, &dtop, &dbot));
            std::vector<int64_t> goff((size_t)nsub + 1), grid((size_t)s5[3] + 1), m7((size_t)nmap * 7 + 1),
                c4((size_t)ncopy * 4 + 1);
            sps_detail::check(sps_fdmt_get_plan_iter(m_h.get(), i, nullptr, goff.data(), grid.data(), nullptr,
                                                     m7.data(), nullptr, c4.data(), nullptr, nullptr));
            p.df_top.push_back(dtop);
            p.df_bot.push_back(dbot);
            std::vector<DtGridType> grids;
            std::vector<SizeType> sub_idx;
            SizeType off = 0;
            for (int64_t s = 0; s < nsub; ++s) {
                DtGridType g;
                for (int64_t k = goff[(size_t)s]; k < goff[(size_t)s + 1]; ++k) g.push_back((SizeType)grid[(size_t)k]);
                sub_idx.push_back(off);
                off += g.size() * m_nsamps;
                grids.push_back(std::move(g));
            }
            std::vector<FDMTCoordType> coords, ccoords;
            std::vector<FDMTCoordMapping> maps, cmaps;
            for (int64_t k = 0; k < nmap; ++k) {
                const int64_t* r = &m7[(size_t)k * 7];
                coords.push_back({(SizeType)r[0], (SizeType)r[1]});
                maps.push_back({{(SizeType)r[2], (SizeType)r[3]}, {(SizeType)r[4], (SizeType)r[5]}, (SizeType)r[6]});
            }
            for (int64_t k = 0; k < ncopy; ++k) {
                const int64_t* r = &c4[(size_t)k * 4];
                ccoords.push_back({(SizeType)r[0], (SizeType)r[1]});
                cmaps.push_back({{SIZE_MAX, SIZE_MAX}, {(SizeType)r[2], (SizeType)r[3]}, 0});
            }
            if (i == 0 || ncopy == 0) top = grids.back();
            p.dt_grid.push_back(std::move(grids));
            p.state_sub_idx.push_back(std::move(sub_idx));
            p.coordinates.push_back(std::move(coords));
            p.coordinates_to_copy.push_back(std::move(ccoords));
            p.mappings.push_back(std::move(maps));
            p.mappings_to_copy.push_back(std::move(cmaps));
            p.dt_grid_sub_top.push_back(top);
        }
    }
};

// FDMTGPU (fdmt_gpu.cuh:68-89): device waterfall, device (or managed) dmt.
class FDMTGPU : public FDMT {
public:
    FDMTGPU(float f_min, float f_max, size_t nchans, size_t nsamps, float tsamp, size_t dt_max, size_t dt_step = 1,
            size_t dt_min = 0, int init_mode = SPS_INIT_MEAN)
        : FDMT(f_min, f_max, nchans, nsamps, tsamp, dt_max, dt_step, dt_min, init_mode) {}

    void execute(const float* waterfall, size_t waterfall_size, float* dmt, size_t dmt_size) override {
        check_inputs(waterfall_size, dmt_size);
        sps_detail::check(sps_fdmt_execute(handle(), waterfall, SPS_F32, waterfall_size, dmt, dmt_size, 0, nullptr));
    }
    // uint8 waterfall promoted to fp32 inside the init kernel
    void execute(const unsigned char* waterfall, size_t waterfall_size, float* dmt, size_t dmt_size) {
        check_inputs(waterfall_size, dmt_size);
        sps_detail::check(sps_fdmt_execute(handle(), waterfall, SPS_U8, waterfall_size, dmt, dmt_size, 0, nullptr));
    }
    void initialise(const float* waterfall, float* state) override {
        const auto& s = get_plan().state_shape[0];
        sps_detail::check(sps_fdmt_initialise(handle(), waterfall, SPS_F32, state, s[3] * s[4], nullptr));
    }
};

// ---------------------------------------------------------------------------
// legacy CFdmt (FDMT/FdmtB.h:6-59, FDMT/FdmtGpu.cuh:5-57)
// ---------------------------------------------------------------------------
typedef float fdmt_type_;  // Constants.h:16

class CFdmtB {
public:
    CFdmtB() : m_nchan(0), m_cols(0), m_Fmin(0), m_Fmax(0), m_imaxDt(0) {}
    CFdmtB(const float Fmin, const float Fmax, int nchan, const int cols, int imaxDt)
        : m_nchan(nchan), m_cols(cols), m_Fmin(Fmin), m_Fmax(Fmax), m_imaxDt(imaxDt) {}
    virtual ~CFdmtB() = default;
    int m_nchan;
    int m_cols;
    float m_Fmin;
    float m_Fmax;
    int m_imaxDt;
    virtual void process_image(fdmt_type_* in, fdmt_type_* out, const bool b_ones) {
        (void)in;
        (void)out;
        (void)b_ones;
    }
    // FdmtB.cpp:144-156
    static unsigned int calc_MaxDT(const float val_fmin_MHz, const float val_fmax_MHz, const float length_of_pulse,
                                   const float val_DM_Max, const int nchan) {
        (void)nchan;
        float t0 = 4148.8 * (1.0 / (val_fmin_MHz * val_fmin_MHz) - 1.0 / (val_fmax_MHz * val_fmax_MHz));
        (void)t0;
        float td = 4148.8 * val_DM_Max * (1.0 / (val_fmin_MHz * val_fmin_MHz) - 1.0 / (val_fmax_MHz * val_fmax_MHz));
        float t_tel = 1.0E-6 / (val_fmax_MHz - val_fmin_MHz);
        float temp = length_of_pulse / t_tel;
        float val_M = ((td / t_tel) / (temp * temp));
        unsigned int r = (unsigned int)(td / val_M / length_of_pulse);
        return (0 == r) ? 1 : r;
    }
    size_t calc_size_input() const { return (size_t)m_cols * m_nchan * sizeof(fdmt_type_); }
    size_t calc_size_output() const { return (size_t)m_cols * m_imaxDt * sizeof(fdmt_type_); }
};

// CFdmtGpu: device in/out; process_image(nullptr, out, true) gives the
// transform of an all-ones input (FDMT/ReadMe.txt:21-27).
class CFdmtGpu : public CFdmtB {
public:
    CFdmtGpu() = default;
    CFdmtGpu(const float Fmin, const float Fmax, int nchan, const int cols, int imaxDt)
        : CFdmtB(Fmin, Fmax, nchan, cols, imaxDt) {
        create();
    }
    CFdmtGpu(const CFdmtGpu& r) : CFdmtB(r) {
        if (r.m_h) create();
    }
    CFdmtGpu& operator=(const CFdmtGpu& r) {
        if (this == &r) return *this;
        CFdmtB::operator=(r);
        m_h.reset();
        if (r.m_h) create();
        return *this;
    }
    void process_image(fdmt_type_* d_in, fdmt_type_* d_out, const bool b_ones) override {
        if (!m_h) throw std::runtime_error("CFdmtGpu is not configured");
        sps_detail::check(sps_fdmt_execute(m_h.get(), d_in, SPS_F32, (size_t)m_nchan * m_cols, d_out,
                                           (size_t)m_imaxDt * m_cols, b_ones ? 1 : 0, nullptr));
    }
    int m_iNumIter() const {
        int64_t n = 0;
        sps_detail::check(sps_fdmt_get_niters(m_h.get(), &n));
        return (int)n;
    }

private:
    std::shared_ptr<sps_fdmt> m_h;
    void create() {
        sps_fdmt* h = nullptr;
        sps_detail::check(sps_cfdmt_create(m_Fmin, m_Fmax, m_nchan, m_cols, m_imaxDt, &h));
        m_h.reset(h, [](sps_fdmt* p) { sps_fdmt_destroy(p); });
    }
};
