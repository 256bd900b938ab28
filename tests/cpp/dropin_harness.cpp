// dropin_harness.cpp -- exercises the drop-in C++ headers the way the
// reference harnesses use the original classes:
//   * fdmt_pr_restr/kernel.cu:355-427: an FDMTCPU sizes the output, then the
//     transform runs through an FDMT* that is an FDMTGPU (device input,
//     managed output) or an FDMTCPU (host input and output);
//   * fdmt_class_v0/kernel.cu:94-136 and CDMT/Chunk_gpu.cu:175-181: CFdmtGpu
//     incl. the ones transform and copy-assignment, and CFdmtCpu on host data;
//   * clusterization_forpravir/clusterization_test.cu:60-270: gathers,
//     clustering and the delegate log on the 3 x 16 test image;
//   * select_delegates/select_delegates.cu (thrust::device_vector forms);
//   * statistical_preprocessing::clean_fdmtInput_and_create_normInput;
//   * detection/kernel.cu:645-726: doDetectSignal_v1 / _v2 / _v2_2 on that
//     harness's input.
// Writes raw outputs for the test to compare.   usage: dropin_harness <outdir>
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "sps/coherent.hpp"
#include "sps/fdmt.hpp"
#include "sps/search.hpp"
#include "sps/thrust_vectors.hpp"
#include "sps/unpack.hpp"

static void dump(const std::string& fn, const void* h, size_t bytes) {
    FILE* f = fopen(fn.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + fn);
    fwrite(h, 1, bytes, f);
    fclose(f);
}

#define HIP_OK(x)                                                                       \
    do {                                                                                \
        if ((x) != hipSuccess) throw std::runtime_error(std::string("HIP call failed: ") + #x); \
    } while (0)

// the call sequence of fdmt_pr_restr/kernel.cu:355-427 (cuda* -> hip*)
static void fdmt_pr_restr_sequence(const std::string& out, const float* h_parrImage, int iImRows, int iImCols,
                                   int iMaxDT, bool gpu, const char* name) {
    const float val_fmin = 1200.0f, val_fmax = 1600.0f;
    fdmt_type_* u_parrImage = NULL;
    fdmt_type_* u_parrImOut = NULL;
    float tsamp = 1.0;
    size_t dt_step = 1;
    size_t dt_min = 0;
    FDMT* pfdmt_cpu = new FDMTCPU(val_fmin, val_fmax, iImRows, iImCols, tsamp, (iMaxDT - 1), dt_step, dt_min);
    const size_t IOutImageRows = pfdmt_cpu->get_dt_grid_final().size();
    delete pfdmt_cpu;
    FDMT* pfdmt = nullptr;
    size_t dmt_size = IOutImageRows * iImCols;
    bool managed = false;
    if (gpu) {
        int mm = 0;
        HIP_OK(hipDeviceGetAttribute(&mm, hipDeviceAttributeManagedMemory, 0));
        HIP_OK(hipMalloc(&u_parrImage, sizeof(fdmt_type_) * iImRows * iImCols));
        HIP_OK(hipMemcpy(u_parrImage, h_parrImage, sizeof(fdmt_type_) * iImRows * iImCols, hipMemcpyHostToDevice));
        if (mm) {
            HIP_OK(hipMallocManaged(&u_parrImOut, iImCols * IOutImageRows * sizeof(fdmt_type_)));
            managed = true;
        } else {
            HIP_OK(hipMalloc(&u_parrImOut, iImCols * IOutImageRows * sizeof(fdmt_type_)));
        }
        pfdmt = new FDMTGPU(val_fmin, val_fmax, iImRows, iImCols, tsamp, iMaxDT - 1, dt_step, dt_min);
    } else {
        u_parrImage = (fdmt_type_*)malloc(sizeof(fdmt_type_) * iImRows * iImCols);
        memcpy(u_parrImage, h_parrImage, sizeof(fdmt_type_) * iImRows * iImCols);
        u_parrImOut = (fdmt_type_*)malloc(sizeof(fdmt_type_) * IOutImageRows * iImCols);
        pfdmt = new FDMTCPU(val_fmin, val_fmax, iImRows, iImCols, tsamp, iMaxDT - 1, dt_step, dt_min);
    }
    for (int i = 0; i < 3; ++i) pfdmt->execute(u_parrImage, iImRows * iImCols, u_parrImOut, dmt_size);
    bool threw = false;
    try {
        pfdmt->execute(u_parrImage, iImRows * iImCols - 1, u_parrImOut, dmt_size);
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    if (!threw) throw std::runtime_error("size check did not throw");
    delete pfdmt;
    std::vector<fdmt_type_> parrImOut(dmt_size);
    if (gpu) {
        HIP_OK(hipMemcpy(parrImOut.data(), u_parrImOut, dmt_size * sizeof(fdmt_type_), hipMemcpyDeviceToHost));
        HIP_OK(hipFree(u_parrImage));
        HIP_OK(hipFree(u_parrImOut));
    } else {
        memcpy(parrImOut.data(), u_parrImOut, dmt_size * sizeof(fdmt_type_));
        free(u_parrImage);
        free(u_parrImOut);
    }
    dump(out + "/" + name, parrImOut.data(), dmt_size * sizeof(fdmt_type_));
    printf("%s: %zu rows%s\n", name, IOutImageRows, managed ? " (managed output)" : "");
}

int main(int argc, char** argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s outdir\n", argv[0]);
        return 2;
    }
    const std::string out = argv[1];
    try {
        // plan-based FDMT on an integer-valued waterfall
        const int rows = 256, cols = 2048, imax = 256;
        std::vector<float> h(rows * cols);
        for (int i = 0; i < rows * cols; ++i) h[i] = (float)((i * 2654435761u >> 7) % 39) - 14.0f;
        dump(out + "/in.bin", h.data(), h.size() * 4);
        fdmt_pr_restr_sequence(out, h.data(), rows, cols, imax, true, "fdmt.bin");
        fdmt_pr_restr_sequence(out, h.data(), rows, cols, imax, false, "fdmt_cpu.bin");
        FDMTCPU::set_num_threads(8);
        if (FDMTCPU::get_num_threads() != 8) throw std::runtime_error("set_num_threads");
        {
            FDMTCPU f(1200.0f, 1600.0f, rows, cols, 1.0f, imax - 1);
            const auto& s0 = f.get_plan().state_shape[0];
            std::vector<float> st(s0[3] * s0[4]);
            f.initialise(h.data(), st.data());
            dump(out + "/state0_cpu.bin", st.data(), st.size() * 4);
        }

        // legacy CFdmtGpu (device) incl. the ones transform and copy-assignment
        float *d_in = nullptr, *d_out2 = nullptr;
        HIP_OK(hipMalloc(&d_in, h.size() * 4));
        HIP_OK(hipMemcpy(d_in, h.data(), h.size() * 4, hipMemcpyHostToDevice));
        CFdmtGpu cf(1200.0f, 1600.0f, rows, cols, imax);
        HIP_OK(hipMalloc(&d_out2, (size_t)imax * cols * 4));
        cf.process_image(d_in, d_out2, false);
        std::vector<float> o2((size_t)imax * cols);
        HIP_OK(hipMemcpy(o2.data(), d_out2, o2.size() * 4, hipMemcpyDeviceToHost));
        dump(out + "/cfdmt.bin", o2.data(), o2.size() * 4);
        CFdmtGpu cf2;
        cf2 = cf;  // the production caller copy-assigns (Chunk_gpu.cu:175-181)
        cf2.process_image(nullptr, d_out2, true);
        HIP_OK(hipMemcpy(o2.data(), d_out2, o2.size() * 4, hipMemcpyDeviceToHost));
        dump(out + "/cfdmt_ones.bin", o2.data(), o2.size() * 4);
        // CFdmtCpu on host buffers, through a CFdmtB pointer
        CFdmtB* pcpu = new CFdmtCpu(1200.0f, 1600.0f, rows, cols, imax);
        std::vector<float> o3((size_t)imax * cols);
        pcpu->process_image(h.data(), o3.data(), false);
        dump(out + "/cfdmt_cpu.bin", o3.data(), o3.size() * 4);
        {
            // its configuration members, and those of a copy
            CFdmtCpu c2(*static_cast<CFdmtCpu*>(pcpu));
            std::ofstream f(out + "/cfdmt_config.txt");
            f << c2.m_iNumIter << "\n";
            for (int i = 0; i <= c2.m_iNumIter; ++i) {
                f << c2.m_parrQuantMtrx_h[i];
                for (int k = 0; k <= c2.m_parrQuantMtrx_h[i]; ++k) f << " " << c2.m_pparrRowsCumSum_h[i][k];
                for (int k = 0; k <= c2.m_parrQuantMtrx_h[i]; ++k) f << " " << c2.m_pparrFreq_h[i][k];
                f << "\n";
            }
        }
        delete pcpu;

        // clusterization on the 3x16 image of clusterization_test.cu:63-81
        std::vector<float> img(48, 2.0f);
        img[4] = 140.f;
        img[11] = 120.f;
        img[2 * 16 + 4] = 100.f;
        float *d_img = nullptr, *d_thr = nullptr;
        int* d_bins = nullptr;
        HIP_OK(hipMalloc(&d_img, 48 * 4));
        HIP_OK(hipMalloc(&d_thr, 4));
        HIP_OK(hipMalloc(&d_bins, 12));
        HIP_OK(hipMemcpy(d_img, img.data(), 48 * 4, hipMemcpyHostToDevice));
        clusterization::digitize_kernel(d_img, 48);  // as clusterization_test.cu does (integer image: unchanged)
        const float thr = 4.0f;
        const int bins[3] = {2, 1, 2};
        HIP_OK(hipMemcpy(d_thr, &thr, 4, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(d_bins, bins, 12, hipMemcpyHostToDevice));
        bool ok = clusterization::clusterization_main(d_img, 3, 16, *d_thr, 2, d_bins, out + "/candidates.log",
                                                      0.00001f, 0.0f);
        if (!ok) throw std::runtime_error("no candidates");
        Cand* d_c1 = nullptr;
        unsigned int n1 = 0;
        clusterization::gather_candidates_in_dynamicalArray(d_img, 3, 16, *d_thr, 1, &d_c1, &n1);
        clusterization::writeCandDeviceArrayToCSV(d_c1, (int)n1, out + "/Candidates1.csv");
        Cand* d_c2 = nullptr;
        unsigned int n2 = 0;
        clusterization::gather_candidates_in_heap_with_plan(d_img, 3, 16, *d_thr, 2, &d_c2, &n2);
        clusterization::writeCandDeviceArrayToCSV(d_c2, (int)n2, out + "/Candidates2.csv");
        printf("candidates %u %u\n", n1, n2);
        // the fixed-array gather into the caller's array: room for all, then for 2
        {
            unsigned int* d_cap = nullptr;
            unsigned int* d_cnt = nullptr;
            Cand* d_fix = nullptr;
            HIP_OK(hipMalloc(&d_cap, 4));
            HIP_OK(hipMalloc(&d_cnt, 4));
            HIP_OK(hipMalloc(&d_fix, sizeof(Cand) * (n1 + 1)));
            unsigned int caps[2] = {n1 + 1, 2};
            std::ofstream f(out + "/fixed.txt");
            for (unsigned int cap : caps) {
                HIP_OK(hipMemcpy(d_cap, &cap, 4, hipMemcpyHostToDevice));
                clusterization::gather_candidates_in_fixedArray_kernel_v0(d_img, 3, 16, *d_thr, 1, d_cap, d_fix,
                                                                          *d_cnt);
                unsigned int cnt = 0;
                HIP_OK(hipMemcpy(&cnt, d_cnt, 4, hipMemcpyDeviceToHost));
                std::vector<Cand> h(std::min(cnt, cap));
                HIP_OK(hipMemcpy(h.data(), d_fix, sizeof(Cand) * h.size(), hipMemcpyDeviceToHost));
                f << cnt;
                for (const Cand& c : h) f << " " << c.mt << "," << c.mdt << "," << c.mwidth << "," << c.msnr;
                f << "\n";
            }
            HIP_OK(hipFree(d_cap));
            HIP_OK(hipFree(d_cnt));
            HIP_OK(hipFree(d_fix));
        }

        // select_delegates over caller groups (thrust::device_vector forms):
        // the 15-candidate example of select_delegates/select_delegates.cu:
        // 323-347 (groups begin at 0, 3, 8 of a permuted index array)
        {
            const float snr[15] = {1.0f, 2.0f, 3.0f, 0.5f, 2.5f, 1.5f, 4.0f, 3.5f, 0.7f, 1.3f, 5.0f, 3.3f, 2.7f, 4.5f, 3.2f};
            std::vector<Cand> hc(15);
            for (int i = 0; i < 15; ++i) hc[(size_t)i] = Cand{i + 1, i + 2, i + 3, snr[i]};
            thrust::device_vector<Cand> heap(hc.begin(), hc.end());
            std::vector<int> hi = {4, 6, 12, 7, 0, 2, 1, 8, 5, 14, 11, 9, 10, 3, 13};
            thrust::device_vector<int> idx(hi.begin(), hi.end());
            std::vector<int> hb = {0, 3, 8};
            thrust::device_vector<int> begin(hb.begin(), hb.end());
            thrust::device_vector<int> dele;
            select_delegates::select_delegates(heap, idx, begin, dele);
            std::vector<int> hd(dele.size());
            thrust::copy(dele.begin(), dele.end(), hd.begin());
            // the kernel form into its own output
            thrust::device_vector<int> dele2(3, -7);
            select_delegates::select_delegates_kernel(thrust::raw_pointer_cast(heap.data()),
                                                      thrust::raw_pointer_cast(idx.data()), 15,
                                                      thrust::raw_pointer_cast(begin.data()),
                                                      thrust::raw_pointer_cast(dele2.data()), 3);
            std::vector<int> hd2(3);
            thrust::copy(dele2.begin(), dele2.end(), hd2.begin());
            std::ofstream f(out + "/delegates.txt");
            for (int v : hd) f << v << " ";
            f << "\n";
            for (int v : hd2) f << v << " ";
            f << "\n";
            select_delegates::write_log(thrust::raw_pointer_cast(heap.data()), 15, begin, idx, dele,
                                        out + "/delegates.log", 0.5f, 2.0f);
        }
        HIP_OK(hipFree(d_c1));
        HIP_OK(hipFree(d_c2));

        // row cleaning without normalisation: row 5 of 64 has 25x the variance
        // (one outlier among n rows lies at most sqrt(n - 1) sigma from the
        // mean variance, so 4 sigma needs more than 17 rows)
        {
            const int R = 64, N = 512;
            std::vector<float> x((size_t)R * N);
            for (int r = 0; r < R; ++r)
                for (int t = 0; t < N; ++t) {
                    const float v = (float)(((r * 131 + t * 71) % 9) - 4);
                    x[(size_t)r * N + t] = (r == 5 ? 5.0f * v : v) + (float)r;
                }
            dump(out + "/clean_in.bin", x.data(), x.size() * 4);
            float *d_x = nullptr, *d_buff = nullptr;
            bool* d_mask = nullptr;
            HIP_OK(hipMalloc(&d_x, x.size() * 4));
            HIP_OK(hipMalloc(&d_mask, x.size()));
            HIP_OK(hipMalloc(&d_buff, (R + 2) * 4));
            HIP_OK(hipMemcpy(d_x, x.data(), x.size() * 4, hipMemcpyHostToDevice));
            statistical_preprocessing::clean_fdmtInput_and_create_normInput(d_x, R, N, d_mask, d_buff);
            std::vector<float> y(x.size()), b(R + 2);
            std::vector<unsigned char> m(x.size());
            HIP_OK(hipMemcpy(y.data(), d_x, y.size() * 4, hipMemcpyDeviceToHost));
            HIP_OK(hipMemcpy(m.data(), d_mask, m.size(), hipMemcpyDeviceToHost));
            HIP_OK(hipMemcpy(b.data(), d_buff, b.size() * 4, hipMemcpyDeviceToHost));
            dump(out + "/clean_out.bin", y.data(), y.size() * 4);
            dump(out + "/clean_mask.bin", m.data(), m.size());
            dump(out + "/clean_buff.bin", b.data(), b.size() * 4);
            HIP_OK(hipFree(d_x));
            HIP_OK(hipFree(d_mask));
            HIP_OK(hipFree(d_buff));
        }

        // the detectors on the input of detection/kernel.cu:645-676
        {
            const int drows = 256;
            const int dcols = (1 << 20) / drows;
            std::vector<float> arr((size_t)drows * dcols, 1.0f), norm((size_t)drows * dcols, 1.0f);
            arr[3] = 20;
            arr[12] = 40;
            arr[8] = 0;
            arr[9] = 0;
            arr[11] = 0;
            for (int i = 13; i <= 17; ++i) arr[(size_t)i] = 0;
            for (int i = 3; i < 5; ++i) norm[(size_t)dcols + i] = 0.001f;
            dump(out + "/det_arr.bin", arr.data(), arr.size() * 4);
            dump(out + "/det_norm.bin", norm.data(), norm.size() * 4);
            fdmt_type_ *d_arr = nullptr, *d_norm = nullptr;
            HIP_OK(hipMalloc(&d_arr, arr.size() * 4));
            HIP_OK(hipMalloc(&d_norm, norm.size() * 4));
            HIP_OK(hipMemcpy(d_arr, arr.data(), arr.size() * 4, hipMemcpyHostToDevice));
            HIP_OK(hipMemcpy(d_norm, norm.data(), norm.size() * 4, hipMemcpyHostToDevice));
            const int WndWidth = 10;
            float valTresh = 5.;
            structOutDetection* pstructOut = nullptr;
            HIP_OK(hipMalloc(&pstructOut, sizeof(structOutDetection)));
            const dim3 blockSize = dim3(256, 1, 1);
            const dim3 gridSize = dim3((dcols + blockSize.x - 1) / blockSize.x, drows, 1);
            size_t sz = blockSize.x * (sizeof(int) + 2 * sizeof(float)) + sizeof(float) + 2 * sizeof(int);
            float* d_pAuxArray = nullptr;
            int *d_pAuxIntArray = nullptr, *d_pAuxWidthArray = nullptr;
            HIP_OK(hipMalloc(&d_pAuxArray, drows * gridSize.x * sizeof(float)));
            HIP_OK(hipMalloc(&d_pAuxIntArray, drows * gridSize.x * sizeof(int)));
            HIP_OK(hipMalloc(&d_pAuxWidthArray, drows * gridSize.x * sizeof(int)));
            std::ofstream f(out + "/detect.txt");
            structOutDetection o{};
            bool b1 = doDetectSignal_v1(d_arr, d_norm, drows, dcols, WndWidth, valTresh, gridSize, blockSize, sz,
                                        d_pAuxArray, d_pAuxIntArray, d_pAuxWidthArray, pstructOut);
            HIP_OK(hipMemcpy(&o, pstructOut, sizeof(o), hipMemcpyDeviceToHost));
            f << "v1 " << b1 << " " << o.iwidth << " " << o.irow << " " << o.icol << " " << o.snr << "\n";
            bool b2 = doDetectSignal_v2(d_arr, d_norm, drows, dcols, WndWidth, valTresh, gridSize, blockSize, 2 * sz,
                                        d_pAuxArray, d_pAuxIntArray, d_pAuxWidthArray, pstructOut);
            HIP_OK(hipMemcpy(&o, pstructOut, sizeof(o), hipMemcpyDeviceToHost));
            f << "v2 " << b2 << " " << o.iwidth << " " << o.irow << " " << o.icol << " " << o.snr << "\n";
            char* d_pBuff = nullptr;
            HIP_OK(hipMalloc(&d_pBuff, gridSize.x * sizeof(float)));
            bool b3 = doDetectSignal_v2_2(d_arr, d_norm, drows, dcols, WndWidth, valTresh, gridSize, blockSize,
                                          2 * sz, d_pAuxArray, d_pAuxIntArray, d_pAuxWidthArray, pstructOut, d_pBuff);
            HIP_OK(hipMemcpy(&o, pstructOut, sizeof(o), hipMemcpyDeviceToHost));
            f << "v2_2 " << b3 << " " << o.iwidth << " " << o.irow << " " << o.icol << " " << o.snr << "\n";
            HIP_OK(hipFree(d_arr));
            HIP_OK(hipFree(d_norm));
            HIP_OK(hipFree(pstructOut));
            HIP_OK(hipFree(d_pAuxArray));
            HIP_OK(hipFree(d_pAuxIntArray));
            HIP_OK(hipFree(d_pAuxWidthArray));
            HIP_OK(hipFree(d_pBuff));
        }
        HIP_OK(hipFree(d_in));
        HIP_OK(hipFree(d_out2));
        HIP_OK(hipFree(d_img));
        HIP_OK(hipFree(d_thr));
        HIP_OK(hipFree(d_bins));

        // CChunk_gpu::process through CoherentChunk: 8-bit GUPPI voltages
        // (hashed noise with two broadband bursts; the test rebuilds the same
        // bytes) unpacked, then the whole coherent chain and the clustering
        {
            const int npol = 4, nchan = 8, len_sft = 8, nbin = 1024, nfft = 4, nov = 128;
            const int nsamp = (nbin - 2 * nov) * nfft;
            std::vector<signed char> raw((size_t)nchan * nsamp * npol);
            for (size_t i = 0; i < raw.size(); ++i) {
                const int s = (int)((i / npol) % nsamp);
                int v = (int)(((unsigned)i * 2654435761u >> 7) % 41) - 20;
                if (s >= 1500 && s < 1564)
                    v *= 4;
                else if (s >= 400 && s < 432)
                    v *= 2;
                raw[i] = (signed char)v;
            }
            char* d_raw = nullptr;
            float2* d_blocks = nullptr;
            float* d_cimg = nullptr;
            int* d_cbins = nullptr;
            HIP_OK(hipMalloc(&d_raw, raw.size()));
            HIP_OK(hipMemcpy(d_raw, raw.data(), raw.size(), hipMemcpyHostToDevice));
            HIP_OK(hipMalloc(&d_blocks, sizeof(float2) * (size_t)(npol / 2) * nfft * nchan * nbin));
            unpack_chunk_guppi_gpu(nsamp, nov, nbin, nfft, npol, nchan, d_raw, d_blocks);
            CoherentChunk cc(1200.0f, 1600.0f, npol, nchan, len_sft, nbin, nfft, nov, 1e-6f);
            const std::vector<double> dms{0.0, 30.0};
            const size_t nimg = dms.size() * (size_t)len_sft * (size_t)cc.msamp();
            HIP_OK(hipMalloc(&d_cimg, nimg * sizeof(float)));
            const int h_cbins[3] = {2, 1, 2};
            HIP_OK(hipMalloc(&d_cbins, sizeof(h_cbins)));
            HIP_OK(hipMemcpy(d_cbins, h_cbins, sizeof(h_cbins), hipMemcpyHostToDevice));
            const bool found = cc.process(d_blocks, dms, 6.0f, d_cimg, 4, d_cbins, out + "/coherent_candidates.log");
            std::vector<float> cimg(nimg);
            HIP_OK(hipMemcpy(cimg.data(), d_cimg, nimg * sizeof(float), hipMemcpyDeviceToHost));
            dump(out + "/coherent_image.bin", cimg.data(), nimg * sizeof(float));
            std::ofstream(out + "/coherent.txt") << (found ? 1 : 0) << " " << cc.msamp() << "\n";
            HIP_OK(hipFree(d_raw));
            HIP_OK(hipFree(d_blocks));
            HIP_OK(hipFree(d_cimg));
            HIP_OK(hipFree(d_cbins));
        }
    } catch (const std::exception& e) {
        fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    printf("harness ok\n");
    return 0;
}
