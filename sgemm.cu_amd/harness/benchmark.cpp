// benchmark.cpp -- GFLOP/s sweep over square sizes, the HIP counterpart of the
// reference's benchmark.cu (benchmark.cu:22-209, run by benchmark.sh:3-43).
//
//   ./bin/sgemm_benchmark [--mmin=1024] [--mstep=512] [--npts=21] [--wmsize=4096]
//                         [--wniter=500] [--fname=sgemm.cu] [--savedir=benchmark_results]
//                         [--form=f32|bf16x3|rocblas] [--config=ID]
//
// Same protocol as the reference: warm-up of `wniter` calls at wmsize^3, then
// for each size s = mmin + i*mstep (i < npts): n_iter = max(1000 *
// exp(-(s - mmin)/3100), 4) event-timed calls with a cache flush between
// them (outside the timed region), the mean of the second half of the
// iterations, FLOP = 2*s^3, alpha = 1, beta = 0, inputs uniform in [0, 1).
// Writes "<size> <GFLOP/s>" lines to <savedir>/<fname>.txt, the file format
// the reference's plot script reads (plot_benchmark_results.py:20-30).
// --form=rocblas times rocblas_sgemm on the same inputs and names the file
// rocBLAS.txt, as the reference's CUBLAS=1 build names cuBLAS.txt
// (benchmark.cu:35-45,150-164).  --config=ID runs kernel config ID of the
// form (sgemm_run_config) instead of the one the library picks.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <string>
#include <vector>

#include "sgemm.h"
#include "harness_util.hpp"

using namespace harness;

namespace {

struct Runner {
    std::string form;
    int config;
    rocblas_handle handle = nullptr;

    Runner(const std::string& f, int cfg) : form(f), config(cfg) {
        if (form == "rocblas" && rocblas_create_handle(&handle) != rocblas_status_success) {
            std::fprintf(stderr, "rocblas_create_handle failed\n");
            std::exit(EXIT_FAILURE);
        }
    }
    ~Runner() {
        if (handle) rocblas_destroy_handle(handle);
    }
    void operator()(int s, const float* alpha, float* A, float* B, const float* beta, float* C) {
        int st = SGEMM_STATUS_SUCCESS;
        if (config >= 0 && form != "rocblas") {
            st = sgemm_run_config(form == "bf16x3" ? SGEMM_FORM_BF16X3 : SGEMM_FORM_F32, config, s,
                                  s, s, alpha, A, s, B, s, beta, C, s, nullptr);
        } else if (form == "bf16x3") {
            st = sgemm_bf16x3(s, s, s, alpha, A, s, B, s, beta, C, s, nullptr);
        } else if (form == "rocblas") {
            if (rocblas_sgemm(handle, rocblas_operation_none, rocblas_operation_none, s, s, s,
                              alpha, A, s, B, s, beta, C, s) != rocblas_status_success)
                st = SGEMM_STATUS_HIP_ERROR;
        } else {
            sgemm(s, s, s, alpha, A, s, B, s, beta, C, s);
            st = sgemm_get_last_error();
        }
        if (st != SGEMM_STATUS_SUCCESS) {
            std::fprintf(stderr, "GEMM failed at size %d: %s (%s)\n", s, sgemm_status_string(st),
                         sgemm_hip_error_string());
            std::exit(EXIT_FAILURE);
        }
    }
};

}  // namespace

int main(int argc, char** argv) {
    const int mmin = flag_int(argc, argv, "mmin", 1024);
    const int mstep = flag_int(argc, argv, "mstep", 512);
    const int npts = flag_int(argc, argv, "npts", 21);
    const int wmsize = flag_int(argc, argv, "wmsize", 4096);
    const int wniter = flag_int(argc, argv, "wniter", 500);
    const std::string savedir = flag(argc, argv, "savedir", "benchmark_results");
    const std::string form = flag(argc, argv, "form", "f32");
    std::string fname = flag(argc, argv, "fname", "sgemm.cu");
    if (form == "rocblas") fname = "rocBLAS";

    const int config = flag_int(argc, argv, "config", -1);
    Runner run(form, config);
    CacheFlusher flusher;
    const float alpha = 1.0f, beta = 0.0f;

    if (wniter > 0) {
        std::printf("=========================\nWarm-up\n=========================\n");
        const size_t n2 = (size_t)wmsize * wmsize;
        std::vector<float> h(n2);
        float *A, *B, *C;
        HIP_CHECK(hipMalloc(&A, n2 * 4));
        HIP_CHECK(hipMalloc(&B, n2 * 4));
        HIP_CHECK(hipMalloc(&C, n2 * 4));
        init_unit(h, 11);
        HIP_CHECK(hipMemcpy(A, h.data(), n2 * 4, hipMemcpyHostToDevice));
        init_unit(h, 12);
        HIP_CHECK(hipMemcpy(B, h.data(), n2 * 4, hipMemcpyHostToDevice));
        for (int i = 0; i < wniter; ++i) {
            run(wmsize, &alpha, A, B, &beta, C);
            HIP_CHECK(hipDeviceSynchronize());
        }
        std::printf("%d / %d\n\n", wniter, wniter);
        HIP_CHECK(hipFree(A));
        HIP_CHECK(hipFree(B));
        HIP_CHECK(hipFree(C));
    }

    std::printf("=========================\nBenchmark (%s)\n=========================\n",
                form.c_str());
    std::vector<int> sizes, gflops;
    hipEvent_t start, stop;
    HIP_CHECK(hipEventCreate(&start));
    HIP_CHECK(hipEventCreate(&stop));
    for (int i = 0; i < npts; ++i) {
        const int s = mmin + i * mstep;
        const size_t n2 = (size_t)s * s;
        std::vector<float> h(n2);
        float *A, *B, *C;
        HIP_CHECK(hipMalloc(&A, n2 * 4));
        HIP_CHECK(hipMalloc(&B, n2 * 4));
        HIP_CHECK(hipMalloc(&C, n2 * 4));
        init_unit(h, 2 * (uint64_t)s + 1);
        HIP_CHECK(hipMemcpy(A, h.data(), n2 * 4, hipMemcpyHostToDevice));
        init_unit(h, 2 * (uint64_t)s + 2);
        HIP_CHECK(hipMemcpy(B, h.data(), n2 * 4, hipMemcpyHostToDevice));
        const double gflop = 2.0 * s * (double)s * s * 1e-9;
        const int n_iter = std::max((int)(1000 * std::exp((-s + mmin) / 3100.0)), 4);
        std::vector<float> ms(n_iter);
        for (int j = 0; j < n_iter; ++j) {
            HIP_CHECK(hipEventRecord(start, nullptr));
            run(s, &alpha, A, B, &beta, C);
            HIP_CHECK(hipEventRecord(stop, nullptr));
            HIP_CHECK(hipEventSynchronize(stop));
            HIP_CHECK(hipEventElapsedTime(&ms[j], start, stop));
            flusher.flush();
        }
        double avg = 0;
        const int mid = n_iter / 2;
        for (int j = mid; j < n_iter; ++j) avg += ms[j];
        avg /= (n_iter - mid);
        sizes.push_back(s);
        gflops.push_back((int)(gflop / (avg * 1e-3)));
        std::printf("m=n=k=%d:\n%s %8d GFLOP/s\n\n", s, fname.c_str(), gflops.back());
        HIP_CHECK(hipFree(A));
        HIP_CHECK(hipFree(B));
        HIP_CHECK(hipFree(C));
    }
    HIP_CHECK(hipEventDestroy(start));
    HIP_CHECK(hipEventDestroy(stop));
    std::filesystem::create_directories(savedir);
    const std::string path = savedir + "/" + fname + ".txt";
    std::ofstream f(path);
    for (size_t i = 0; i < sizes.size(); ++i) f << sizes[i] << " " << gflops[i] << "\n";
    std::printf("Benchmark data stored in %s\n=========================\n", path.c_str());
    return EXIT_SUCCESS;
}
