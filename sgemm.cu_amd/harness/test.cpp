// test.cpp -- correctness sweep over square sizes, the HIP counterpart of the
// reference's test.cu (test.cu:18-121, run by test.sh:3-13).
//
//   ./bin/sgemm_test [--mmin=2] [--mmax=1024] [--mstep=1] [--savedir=test_results]
//                    [--form=f32|bf16x3]
//
// For every size s in mmin..mmax step mstep: A, B, C uniform in [0, 1)
// (seeded), alpha = 1.5, beta = 0.5, m = n = k = s, column-major with
// ld = s; the library's sgemm is compared with its naive sgemm_basic by the
// reference's rule (relative error > 1e-4 fails, NaN/Inf fail).  Writes
// <savedir>/sgemm.cu.txt like the reference.  Unlike the reference
// (test.cu:120) the exit status is non-zero when any size fails.
#include <hip/hip_runtime.h>

#include <filesystem>
#include <fstream>
#include <string>
#include <vector>

#include "sgemm.h"
#include "harness_util.hpp"

using namespace harness;

int main(int argc, char** argv) {
    const int mmin = flag_int(argc, argv, "mmin", 2);
    const int mmax = flag_int(argc, argv, "mmax", 1024);
    const int mstep = std::max(1, flag_int(argc, argv, "mstep", 1));
    const std::string savedir = flag(argc, argv, "savedir", "test_results");
    const std::string form = flag(argc, argv, "form", "f32");
    const bool split = form == "bf16x3";

    std::printf("=========================\nTesting (%s)...\n=========================\n",
                form.c_str());
    std::string full, failed;
    int n_tests = 0, n_failed = 0;
    const float alpha = 1.5f, beta = 0.5f;
    for (int s = mmin; s <= mmax; s += mstep, ++n_tests) {
        const size_t n2 = (size_t)s * s;
        std::vector<float> A(n2), B(n2), C(n2), out(n2), ref(n2);
        init_unit(A, 3 * (uint64_t)s + 1);
        init_unit(B, 3 * (uint64_t)s + 2);
        init_unit(C, 3 * (uint64_t)s + 3);
        float *dA, *dB, *dC, *dR;
        HIP_CHECK(hipMalloc(&dA, n2 * 4));
        HIP_CHECK(hipMalloc(&dB, n2 * 4));
        HIP_CHECK(hipMalloc(&dC, n2 * 4));
        HIP_CHECK(hipMalloc(&dR, n2 * 4));
        HIP_CHECK(hipMemcpy(dA, A.data(), n2 * 4, hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(dB, B.data(), n2 * 4, hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(dC, C.data(), n2 * 4, hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(dR, C.data(), n2 * 4, hipMemcpyHostToDevice));
        int st;
        if (split) {
            st = sgemm_bf16x3(s, s, s, &alpha, dA, s, dB, s, &beta, dC, s, nullptr);
        } else {
            sgemm(s, s, s, &alpha, dA, s, dB, s, &beta, dC, s);
            st = sgemm_get_last_error();
        }
        sgemm_basic(s, s, s, &alpha, dA, s, dB, s, &beta, dR, s);
        const int st2 = sgemm_get_last_error();
        HIP_CHECK(hipMemcpy(out.data(), dC, n2 * 4, hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(ref.data(), dR, n2 * 4, hipMemcpyDeviceToHost));
        CompareResult r = compare(ref.data(), out.data(), n2, 1e-4);
        const bool ok = r.equal && st == SGEMM_STATUS_SUCCESS && st2 == SGEMM_STATUS_SUCCESS;
        if (!ok) {
            ++n_failed;
            failed += std::to_string(n_tests) + " ";
        }
        std::printf("Test #%d | matsize = %d | %s\n", n_tests, s, ok ? "PASSED" : "FAILED");
        full += "Test #" + std::to_string(n_tests) + " Matsize = " + std::to_string(s) +
                ": n_false = " + std::to_string(r.n_false) + ", n_nan = " +
                std::to_string(r.n_nan) + ", n_inf = " + std::to_string(r.n_inf) + "\n";
        HIP_CHECK(hipFree(dA));
        HIP_CHECK(hipFree(dB));
        HIP_CHECK(hipFree(dC));
        HIP_CHECK(hipFree(dR));
    }
    std::string summary = "\n=============== SUMMARY ===============\n";
    summary += "PASSED: " + std::to_string(n_tests - n_failed) + " / " + std::to_string(n_tests) +
               "\n";
    summary += "FAILED: " + (failed.empty() ? std::string("0") : failed) + "\n";
    std::printf("%s", summary.c_str());
    std::filesystem::create_directories(savedir);
    const std::string path = savedir + "/sgemm.cu.txt";
    std::ofstream f(path);
    f << full << summary;
    std::printf("Test results stored in %s\n", path.c_str());
    return n_failed ? EXIT_FAILURE : EXIT_SUCCESS;
}
