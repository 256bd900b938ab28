// ref_shim.cpp -- extern "C" entry points over the reference's own CPU FDMT
// classes, compiled together with the reference sources where they lie
// (see oracle/Makefile).  TEST INFRASTRUCTURE ONLY: used to pin the oracle's
// restatement; never part of the product.
//
// Built three times into oracle/_ref/:
//   REF_PRAVIR   -> FDMT_PRAVIR/{fdmtUtils,fdmt_base,fdmt_cpu}.cpp
//   REF_RESTRICT -> FDMT_PRAVIR_RESTRICT/{fdmtUtils,fdmt_base,fdmt_cpu}.cpp
//   REF_CFDMT    -> FDMT/{FdmtB,FdmtCpu}.cpp
#include <cstdint>
#include <cstring>
#include <vector>

#if defined(REF_PRAVIR) || defined(REF_RESTRICT)
#include <fdmt_cpu.hpp>

extern "C" {

// FDMTCPU(f_min, f_max, nchans, nsamps, tsamp, dt_max, dt_step, dt_min)
// (fdmt_cpu.hpp:5-20); returns nullptr when the plan builder throws.
void* ref_fdmt_create(float fmin, float fmax, int64_t nchans, int64_t nsamps, float tsamp,
                      int64_t dt_max, int64_t dt_step, int64_t dt_min) {
    try {
        return new FDMTCPU(fmin, fmax, (size_t)nchans, (size_t)nsamps, tsamp, (size_t)dt_max,
                           (size_t)dt_step, (size_t)dt_min);
    } catch (...) {
        return nullptr;
    }
}
void ref_fdmt_destroy(void* h) { delete static_cast<FDMTCPU*>(h); }
// FDMTCPU::set_num_threads (fdmt_cpu.hpp:9)
void ref_fdmt_set_num_threads(int n) { FDMTCPU::set_num_threads(n); }

int64_t ref_fdmt_niters(void* h) { return (int64_t) static_cast<FDMTCPU*>(h)->get_niters(); }

void ref_fdmt_state_shape(void* h, int64_t i, int64_t* out5) {
    const auto& s = static_cast<FDMTCPU*>(h)->get_plan().state_shape[(size_t)i];
    for (int k = 0; k < 5; ++k) out5[k] = (int64_t)s[k];
}

int64_t ref_fdmt_nsub(void* h, int64_t i) {
    return (int64_t) static_cast<FDMTCPU*>(h)->get_plan().dt_grid[(size_t)i].size();
}

// concatenated dt grids with offsets (nsub+1)
void ref_fdmt_grid(void* h, int64_t i, int64_t* off, int64_t* grid) {
    const auto& g = static_cast<FDMTCPU*>(h)->get_plan().dt_grid[(size_t)i];
    int64_t n = 0;
    off[0] = 0;
    for (size_t s = 0; s < g.size(); ++s) {
        for (auto v : g[s]) grid[n++] = (int64_t)v;
        off[s + 1] = n;
    }
}

int64_t ref_fdmt_nmap(void* h, int64_t i) {
    return (int64_t) static_cast<FDMTCPU*>(h)->get_plan().mappings[(size_t)i].size();
}
int64_t ref_fdmt_ncopy(void* h, int64_t i) {
    return (int64_t) static_cast<FDMTCPU*>(h)->get_plan().mappings_to_copy[(size_t)i].size();
}

// map7 rows: out_sub, out_dt, head_sub, head_dt, tail_sub, tail_dt, offset
// copy4 rows: out_sub, out_dt, tail_sub, tail_dt
void ref_fdmt_maps(void* h, int64_t i, int64_t* map7, int64_t* copy4) {
    const auto& p = static_cast<FDMTCPU*>(h)->get_plan();
    const auto& c = p.coordinates[(size_t)i];
    const auto& m = p.mappings[(size_t)i];
    for (size_t k = 0; k < m.size(); ++k) {
        int64_t* r = map7 + 7 * k;
        r[0] = (int64_t)c[k].first; r[1] = (int64_t)c[k].second;
        r[2] = (int64_t)m[k].head.first; r[3] = (int64_t)m[k].head.second;
        r[4] = (int64_t)m[k].tail.first; r[5] = (int64_t)m[k].tail.second;
        r[6] = (int64_t)m[k].offset;
    }
    const auto& cc = p.coordinates_to_copy[(size_t)i];
    const auto& mc = p.mappings_to_copy[(size_t)i];
    for (size_t k = 0; k < mc.size(); ++k) {
        int64_t* r = copy4 + 4 * k;
        r[0] = (int64_t)cc[k].first; r[1] = (int64_t)cc[k].second;
        r[2] = (int64_t)mc[k].tail.first; r[3] = (int64_t)mc[k].tail.second;
    }
}

void ref_fdmt_df(void* h, int64_t i, float* df_top, float* df_bot) {
    const auto& p = static_cast<FDMTCPU*>(h)->get_plan();
    *df_top = p.df_top[(size_t)i];
    *df_bot = p.df_bot[(size_t)i];
}

// FDMTCPU::execute; returns 0, or -1 when check_inputs / the kernels throw.
int ref_fdmt_execute(void* h, const float* x, int64_t x_size, float* dmt, int64_t dmt_size) {
    try {
        static_cast<FDMTCPU*>(h)->execute(x, (size_t)x_size, dmt, (size_t)dmt_size);
        return 0;
    } catch (...) {
        return -1;
    }
}

void ref_fdmt_dm_grid(void* h, float* out) {
    auto v = static_cast<FDMTCPU*>(h)->get_dm_grid_final();
    std::memcpy(out, v.data(), v.size() * sizeof(float));
}

}  // extern "C"
#endif

#ifdef REF_CFDMT
#include "FdmtCpu.h"

extern "C" {

void* ref_cfdmt_create(float fmin, float fmax, int nchan, int cols, int imaxdt) {
    return new CFdmtCpu(fmin, fmax, nchan, cols, imaxdt);
}
void ref_cfdmt_destroy(void* h) { delete static_cast<CFdmtCpu*>(h); }
int ref_cfdmt_niter(void* h) { return static_cast<CFdmtCpu*>(h)->m_iNumIter; }
int ref_cfdmt_quant(void* h, int i) { return static_cast<CFdmtCpu*>(h)->m_parrQuantMtrx_h[i]; }
void ref_cfdmt_config(void* h, int i, int* cumsum, float* freq) {
    auto* f = static_cast<CFdmtCpu*>(h);
    const int q = f->m_parrQuantMtrx_h[i];
    std::memcpy(cumsum, f->m_pparrRowsCumSum_h[i], sizeof(int) * (size_t)(q + 1));
    std::memcpy(freq, f->m_pparrFreq_h[i], sizeof(float) * (size_t)(q + 1));
}
// CFdmtB::calc_MaxDT (FdmtB.h:33-34, FdmtB.cpp:144-156)
unsigned int ref_calc_maxdt(float fmin, float fmax, float length_of_pulse, float dm_max, int nchan) {
    return CFdmtB::calc_MaxDT(fmin, fmax, length_of_pulse, dm_max, nchan);
}
// CFdmtCpu::process_image(in, out, b_ones); in may be NULL with b_ones.
void ref_cfdmt_process(void* h, float* x, float* out, int b_ones) {
    static_cast<CFdmtCpu*>(h)->process_image(x, out, b_ones != 0);
}

}  // extern "C"
#endif
