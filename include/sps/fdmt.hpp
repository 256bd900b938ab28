// sps/fdmt.hpp -- drop-in C++ FDMT classes over the libsps C ABI.
//
// Mirrors FDMT_PRAVIR_RESTRICT/fdmt_base.hpp:8-85 (FDMT, FDMTPlan and its
// types), fdmt_cpu.hpp:5-20 (FDMTCPU), fdmt_gpu.cuh:68-89 (FDMTGPU) and
// FDMT/FdmtB.h:6-59, FdmtCpu.h:6-34, FdmtGpu.cuh:5-57 (CFdmtB, CFdmtCpu,
// CFdmtGpu).  FDMTGPU and CFdmtGpu take HIP device pointers and issue their
// work on the default stream like the reference; FDMTCPU and CFdmtCpu take
// host pointers and return when the result is in host memory (the transform
// runs on the GPU either way).  Errors throw the reference's exception types:
// std::invalid_argument for size/argument errors, std::runtime_error otherwise.
#pragma once

#include <array>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
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

// FDMTCPU (FDMT_PRAVIR_RESTRICT/fdmt_cpu.hpp:5-20): host waterfall and host
// dmt.  The transform itself runs on the GPU (copy in, transform, copy back,
// synchronous); set_num_threads records the thread count a caller asks for,
// which no host computation here uses.
class FDMTCPU : public FDMT {
public:
    FDMTCPU(float f_min, float f_max, size_t nchans, size_t nsamps, float tsamp, size_t dt_max, size_t dt_step = 1,
            size_t dt_min = 0, int init_mode = SPS_INIT_MEAN)
        : FDMT(f_min, f_max, nchans, nsamps, tsamp, dt_max, dt_step, dt_min, init_mode) {}

    static void set_num_threads(int nthreads) { threads_ref() = nthreads; }
    static int get_num_threads() { return threads_ref(); }

    void execute(const float* waterfall, size_t waterfall_size, float* dmt, size_t dmt_size) override {
        check_inputs(waterfall_size, dmt_size);
        sps_detail::check(sps_fdmt_execute_host(handle(), waterfall, waterfall_size, dmt, dmt_size, 0));
    }
    void initialise(const float* waterfall, float* state) override {
        const auto& s = get_plan().state_shape[0];
        sps_detail::check(sps_fdmt_initialise_host(handle(), waterfall, state, s[3] * s[4]));
    }

private:
    static int& threads_ref() {
        static int n = 0;
        return n;
    }
};

// ---------------------------------------------------------------------------
// legacy CFdmt (FDMT/FdmtB.h:6-59, FDMT/FdmtCpu.h:6-34, FDMT/FdmtGpu.cuh:5-57)
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
    // FdmtB.cpp:144-156.  The dispersion delay across the band of a pulse at
    // val_DM_Max is td = 4148.8 * DM * (fmin^-2 - fmax^-2) seconds and one
    // sample of a critically sampled band lasts ts = 1e-6 / (fmax - fmin)
    // seconds.  With the pulse length L (seconds) the result is
    //   td / (td / ts / (L / ts)^2) / L = L / ts
    // evaluated in that order in fp32, truncated, and at least 1.
    static unsigned int calc_MaxDT(const float val_fmin_MHz, const float val_fmax_MHz, const float length_of_pulse,
                                   const float val_DM_Max, const int nchan) {
        (void)nchan;
        const double band = 1.0 / (val_fmin_MHz * val_fmin_MHz) - 1.0 / (val_fmax_MHz * val_fmax_MHz);
        const float delay_s = (float)(4148.8 * val_DM_Max * band);
        const float sample_s = (float)(1.0E-6 / (val_fmax_MHz - val_fmin_MHz));
        const float pulse_samples = length_of_pulse / sample_s;
        const float ratio = (delay_s / sample_s) / (pulse_samples * pulse_samples);
        const unsigned int n = (unsigned int)(delay_s / ratio / length_of_pulse);
        return n == 0 ? 1u : n;
    }
    size_t calc_size_input() const { return (size_t)m_cols * m_nchan * sizeof(fdmt_type_); }
    size_t calc_size_output() const { return (size_t)m_cols * m_imaxDt * sizeof(fdmt_type_); }
};

// CFdmtCpu (FDMT/FdmtCpu.h:6-34): host input and output.  The public
// configuration members are those of the reference (malloc'ed tables, one
// entry per iteration 0..m_iNumIter), deep-copied on copy.  process_image
// runs the transform on the GPU (copy in, transform, copy back, synchronous);
// process_image(nullptr, out, true) gives the transform of an all-ones input
// (FDMT/ReadMe.txt:21-27).
class CFdmtCpu : public CFdmtB {
public:
    CFdmtCpu() : CFdmtB() {}
    CFdmtCpu(const float Fmin, const float Fmax, int nchan, const int cols, int imaxDt)
        : CFdmtB(Fmin, Fmax, nchan, cols, imaxDt) {
        create_handle();
        create_config(m_pparrRowsCumSum_h, m_pparrFreq_h, &m_parrQuantMtrx_h, &m_iNumIter);
    }
    CFdmtCpu(const CFdmtCpu& r) : CFdmtB(r) { copy_from(r); }
    CFdmtCpu& operator=(const CFdmtCpu& r) {
        if (this == &r) return *this;
        CFdmtB::operator=(r);
        release_config();
        m_h.reset();
        copy_from(r);
        return *this;
    }
    ~CFdmtCpu() override { release_config(); }

    int** m_pparrRowsCumSum_h = nullptr;
    float** m_pparrFreq_h = nullptr;
    int* m_parrQuantMtrx_h = nullptr;
    int m_iNumIter = 0;

    void process_image(fdmt_type_* piarrImgInp, fdmt_type_* piarrImgOut, const bool b_ones) override {
        if (!m_h) throw std::runtime_error("CFdmtCpu is not configured");
        sps_detail::check(sps_fdmt_execute_host(m_h.get(), piarrImgInp, (size_t)m_nchan * m_cols, piarrImgOut,
                                                (size_t)m_imaxDt * m_cols, b_ones ? 1 : 0));
    }

    // FdmtCpu.cpp:143-194: malloc'ed tables, owned by the caller
    virtual void create_config(int**& pparrRowsCumSum, float**& pparrFreq, int** pparrQuantMtrx, int* piNumIter) {
        if (!m_h) throw std::runtime_error("CFdmtCpu is not configured");
        int64_t n = 0;
        sps_detail::check(sps_fdmt_get_niters(m_h.get(), &n));
        *piNumIter = (int)n;
        *pparrQuantMtrx = (int*)malloc(sizeof(int) * (size_t)(n + 1));
        pparrRowsCumSum = (int**)malloc(sizeof(int*) * (size_t)(n + 1));
        pparrFreq = (float**)malloc(sizeof(float*) * (size_t)(n + 1));
        if (!*pparrQuantMtrx || !pparrRowsCumSum || !pparrFreq) throw std::bad_alloc();
        for (int64_t i = 0; i <= n; ++i) {
            int32_t q = 0;
            sps_detail::check(sps_cfdmt_get_config(m_h.get(), i, &q, nullptr, nullptr));
            (*pparrQuantMtrx)[i] = q;
            pparrRowsCumSum[i] = (int*)malloc(sizeof(int) * (size_t)(q + 1));
            pparrFreq[i] = (float*)malloc(sizeof(float) * (size_t)(q + 1));
            if (!pparrRowsCumSum[i] || !pparrFreq[i]) throw std::bad_alloc();
            sps_detail::check(sps_cfdmt_get_config(m_h.get(), i, &q, pparrRowsCumSum[i], pparrFreq[i]));
        }
    }

    sps_fdmt* handle() const { return m_h.get(); }

protected:
    std::shared_ptr<sps_fdmt> m_h;

    void create_handle() {
        sps_fdmt* h = nullptr;
        sps_detail::check(sps_cfdmt_create(m_Fmin, m_Fmax, m_nchan, m_cols, m_imaxDt, &h));
        m_h.reset(h, [](sps_fdmt* p) { sps_fdmt_destroy(p); });
    }

private:
    void release_config() {
        for (int i = 0; i <= m_iNumIter; ++i) {
            if (m_pparrRowsCumSum_h) free(m_pparrRowsCumSum_h[i]);
            if (m_pparrFreq_h) free(m_pparrFreq_h[i]);
        }
        free(m_pparrRowsCumSum_h);
        free(m_pparrFreq_h);
        free(m_parrQuantMtrx_h);
        m_pparrRowsCumSum_h = nullptr;
        m_pparrFreq_h = nullptr;
        m_parrQuantMtrx_h = nullptr;
        m_iNumIter = 0;
    }
    void copy_from(const CFdmtCpu& r) {
        if (!r.m_h) return;
        create_handle();  // the device state is never shared between copies
        m_iNumIter = r.m_iNumIter;
        const size_t n = (size_t)m_iNumIter + 1;
        m_parrQuantMtrx_h = (int*)malloc(sizeof(int) * n);
        m_pparrRowsCumSum_h = (int**)calloc(n, sizeof(int*));
        m_pparrFreq_h = (float**)calloc(n, sizeof(float*));
        if (!m_parrQuantMtrx_h || !m_pparrRowsCumSum_h || !m_pparrFreq_h) throw std::bad_alloc();
        std::memcpy(m_parrQuantMtrx_h, r.m_parrQuantMtrx_h, sizeof(int) * n);
        for (size_t i = 0; i < n; ++i) {
            const size_t q = (size_t)m_parrQuantMtrx_h[i] + 1;
            m_pparrRowsCumSum_h[i] = (int*)malloc(sizeof(int) * q);
            m_pparrFreq_h[i] = (float*)malloc(sizeof(float) * q);
            if (!m_pparrRowsCumSum_h[i] || !m_pparrFreq_h[i]) throw std::bad_alloc();
            std::memcpy(m_pparrRowsCumSum_h[i], r.m_pparrRowsCumSum_h[i], sizeof(int) * q);
            std::memcpy(m_pparrFreq_h[i], r.m_pparrFreq_h[i], sizeof(float) * q);
        }
    }
};

// CFdmtGpu (FDMT/FdmtGpu.cuh:5-57): a CFdmtCpu whose process_image takes
// device input and output and is ordered on the default stream.
class CFdmtGpu : public CFdmtCpu {
public:
    CFdmtGpu() = default;
    CFdmtGpu(const float Fmin, const float Fmax, int nchan, const int cols, int imaxDt)
        : CFdmtCpu(Fmin, Fmax, nchan, cols, imaxDt) {}
    CFdmtGpu(const CFdmtGpu& r) = default;
    CFdmtGpu& operator=(const CFdmtGpu& r) = default;

    void process_image(fdmt_type_* d_in, fdmt_type_* d_out, const bool b_ones) override {
        if (!m_h) throw std::runtime_error("CFdmtGpu is not configured");
        sps_detail::check(sps_fdmt_execute(m_h.get(), d_in, SPS_F32, (size_t)m_nchan * m_cols, d_out,
                                           (size_t)m_imaxDt * m_cols, b_ones ? 1 : 0, nullptr));
    }
};
