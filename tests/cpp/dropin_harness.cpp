// dropin_harness.cpp -- exercises the drop-in C++ headers the way the
// reference harnesses use the original classes (fdmt_pr_restr/kernel.cu:
// 362-424, fdmt_class_v0/kernel.cu:94-136, clusterization_forpravir/
// clusterization_test.cu:60-270).  Writes raw outputs for the test to compare.
//   usage: dropin_harness <outdir>
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>
#include <vector>

#include "sps/fdmt.hpp"
#include "sps/search.hpp"

static void dump(const std::string& fn, const void* h, size_t bytes) {
    FILE* f = fopen(fn.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + fn);
    fwrite(h, 1, bytes, f);
    fclose(f);
}

int main(int argc, char** argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s outdir\n", argv[0]);
        return 2;
    }
    const std::string out = argv[1];
    try {
        // plan-based FDMT on an integer-valued waterfall (fdmt_pr_restr style)
        const int rows = 256, cols = 2048, imax = 256;
        std::vector<float> h(rows * cols);
        for (int i = 0; i < rows * cols; ++i) h[i] = (float)((i * 2654435761u >> 7) % 39) - 14.0f;
        dump(out + "/in.bin", h.data(), h.size() * 4);
        FDMTGPU fdmt(1200.0f, 1600.0f, rows, cols, 1.0f, imax - 1, 1, 0);
        const size_t nout = fdmt.get_dt_grid_final().size() * cols;
        float *d_in = nullptr, *d_out = nullptr;
        hipMalloc(&d_in, h.size() * 4);
        hipMalloc(&d_out, nout * 4);
        hipMemcpy(d_in, h.data(), h.size() * 4, hipMemcpyHostToDevice);
        fdmt.execute(d_in, (size_t)rows * cols, d_out, nout);
        std::vector<float> o(nout);
        hipMemcpy(o.data(), d_out, nout * 4, hipMemcpyDeviceToHost);
        dump(out + "/fdmt.bin", o.data(), nout * 4);
        bool threw = false;
        try {
            fdmt.execute(d_in, (size_t)rows * cols - 1, d_out, nout);
        } catch (const std::invalid_argument&) {
            threw = true;
        }
        if (!threw) throw std::runtime_error("size check did not throw");
        printf("fdmt rows %zu niters %zu df %g\n", fdmt.get_dt_grid_final().size(), fdmt.get_niters(), fdmt.get_df());

        // legacy CFdmtGpu incl. the ones transform (fdmt_class_v0 style)
        CFdmtGpu cf(1200.0f, 1600.0f, rows, cols, imax);
        float* d_out2 = nullptr;
        hipMalloc(&d_out2, (size_t)imax * cols * 4);
        cf.process_image(d_in, d_out2, false);
        std::vector<float> o2((size_t)imax * cols);
        hipMemcpy(o2.data(), d_out2, o2.size() * 4, hipMemcpyDeviceToHost);
        dump(out + "/cfdmt.bin", o2.data(), o2.size() * 4);
        CFdmtGpu cf2;
        cf2 = cf;  // the production caller copy-assigns (Chunk_gpu.cu:175-181)
        cf2.process_image(nullptr, d_out2, true);
        hipMemcpy(o2.data(), d_out2, o2.size() * 4, hipMemcpyDeviceToHost);
        dump(out + "/cfdmt_ones.bin", o2.data(), o2.size() * 4);

        // clusterization on the 3x16 image of clusterization_test.cu:63-81
        std::vector<float> img(48, 2.0f);
        img[4] = 140.f;
        img[11] = 120.f;
        img[2 * 16 + 4] = 100.f;
        float *d_img = nullptr, *d_thr = nullptr;
        int* d_bins = nullptr;
        hipMalloc(&d_img, 48 * 4);
        hipMalloc(&d_thr, 4);
        hipMalloc(&d_bins, 12);
        hipMemcpy(d_img, img.data(), 48 * 4, hipMemcpyHostToDevice);
        clusterization::digitize_kernel(d_img, 48);  // as clusterization_test.cu does (integer image: unchanged)
        const float thr = 4.0f;
        const int bins[3] = {2, 1, 2};
        hipMemcpy(d_thr, &thr, 4, hipMemcpyHostToDevice);
        hipMemcpy(d_bins, bins, 12, hipMemcpyHostToDevice);
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
        hipFree(d_c1);
        hipFree(d_c2);
        hipFree(d_in);
        hipFree(d_out);
        hipFree(d_out2);
        hipFree(d_img);
        hipFree(d_thr);
        hipFree(d_bins);
    } catch (const std::exception& e) {
        fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    printf("harness ok\n");
    return 0;
}
