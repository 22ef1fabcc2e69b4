// autotune.cpp -- tile-shape / split-K autotuner for the grouped GEMM.
//
// No reference counterpart: the reference fixes its tile-group shape at build
// time (apps/sgemm/template.mk:7-9) and sweeps shapes by hand in legacy test
// lists (apps/legacy/gemm/sgemm_512/tests.mk:1-6).  This program measures every
// candidate (tile shape x split-K) per size class on the device and writes the
// tuning file hbs_grouped_plan_create reads:
//   <fp32|bf16x3> <M class> <N class> <K class> <tile id> <split> <ns per work item>
// A size class is the power of two nearest the size in log scale (sizes in
// [c / sqrt2, c * sqrt2) map to c), clamped to [64, 2048].
// Each probe runs a group of identical problems large enough to fill the GPU,
// timed with HIP events over --reps launches after one warm-up launch.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "hbsgemm.h"

#define HIP_CHECK(expr)                                                                  \
    do {                                                                                 \
        hipError_t _e = (expr);                                                          \
        if (_e != hipSuccess) {                                                          \
            fprintf(stderr, "%s: %s\n", #expr, hipGetErrorString(_e));                   \
            exit(2);                                                                     \
        }                                                                                \
    } while (0)

struct Probe {
    int tile, split;
    double ms, ns_item;
};

static std::vector<int> parse_list(const char *s) {
    std::vector<int> v;
    std::string t(s);
    size_t p = 0;
    while (p < t.size()) {
        size_t q = t.find(',', p);
        if (q == std::string::npos) q = t.size();
        v.push_back(atoi(t.substr(p, q - p).c_str()));
        p = q + 1;
    }
    return v;
}

int main(int argc, char **argv) {
    const char *out_path = "tuning.txt";
    std::string prec_arg = "both";
    std::vector<int> classes = {64, 128, 256, 512, 1024, 2048};
    int reps = 5;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--out" && i + 1 < argc) out_path = argv[++i];
        else if (a == "--prec" && i + 1 < argc) prec_arg = argv[++i];
        else if (a == "--classes" && i + 1 < argc) classes = parse_list(argv[++i]);
        else if (a == "--reps" && i + 1 < argc) reps = atoi(argv[++i]);
        else {
            fprintf(stderr,
                    "usage: %s [--out file] [--prec fp32|bf16x3|both] [--classes 64,128,...] "
                    "[--reps r]\n",
                    argv[0]);
            return 2;
        }
    }
    if (reps < 1 || classes.empty()) return 2;
    for (int c : classes)
        if (c < 64 || c > 2048 || (c & (c - 1))) {
            fprintf(stderr, "classes must be powers of two in [64, 2048]\n");
            return 2;
        }
    std::vector<int> precs;
    if (prec_arg == "fp32" || prec_arg == "both") precs.push_back(HBS_PREC_FP32);
    if (prec_arg == "bf16x3" || prec_arg == "both") precs.push_back(HBS_PREC_BF16X3);
    if (precs.empty()) return 2;

    int cus = 256;
    HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
    const int maxc = *std::max_element(classes.begin(), classes.end());
    const size_t elems = (size_t)maxc * maxc;
    // Enough C space for the largest group probed (see `count` below).
    const size_t c_elems = (size_t)8 * cus * 128 * 128 + elems;
    float *A, *B, *C;
    HIP_CHECK(hipMalloc(&A, elems * sizeof(float)));
    HIP_CHECK(hipMalloc(&B, elems * sizeof(float)));
    HIP_CHECK(hipMalloc(&C, c_elems * sizeof(float)));
    if (hbs_fill_random(A, (int64_t)elems, HBS_GEN_UNIFORM, 1, 0, nullptr) ||
        hbs_fill_random(B, (int64_t)elems, HBS_GEN_UNIFORM, 2, 0, nullptr)) {
        fprintf(stderr, "input generation failed\n");
        return 2;
    }
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));

    FILE *out = fopen(out_path, "w");
    if (!out) {
        perror(out_path);
        return 2;
    }
    fprintf(out, "# hbsgemm grouped tuning file: prec Mclass Nclass Kclass tile split ns_per_item\n");
    for (int prec : precs) {
        const char *pname = prec == HBS_PREC_FP32 ? "fp32" : "bf16x3";
        const int ntiles = 4;
        for (int mc : classes) {
            for (int nc : classes) {
                for (int kc : classes) {
                    // identical problems until the largest tile shape (128 x 128) has
                    // >= 8 work items per CU, so every candidate runs on a full GPU
                    const int64_t t128 = (int64_t)((mc + 127) / 128) * ((nc + 127) / 128);
                    const int count = (int)std::max<int64_t>(1, std::min<int64_t>((8 * cus + t128 - 1) / t128, 4096));
                    std::vector<hbs_gemm_problem> probs(count);
                    for (int i = 0; i < count; ++i) {
                        float *Ci = C + (size_t)i * mc * nc;
                        if ((size_t)(i + 1) * mc * nc > c_elems) Ci = C;  // timing only
                        probs[i] = hbs_gemm_problem{A, B, Ci, mc, nc, kc, kc, nc, nc};
                    }
                    Probe best{-1, 0, 1e30, 0.0};
                    for (int tile = 0; tile < ntiles; ++tile) {
                        for (int split : {1, 2, 4}) {
                            if (split > 1 && kc < 512) continue;
                            hbs_grouped_plan *plan = nullptr;
                            int rc = hbs_grouped_plan_create_forced(probs.data(), count, prec, tile,
                                                                    split, &plan);
                            if (rc) {
                                fprintf(stderr, "plan %s %d %d %d tile %d split %d: %s\n", pname,
                                        mc, nc, kc, tile, split, hbs_strerror(rc));
                                return 2;
                            }
                            hbs_grouped_info info;
                            hbs_grouped_plan_info(plan, &info);
                            rc = hbs_grouped_run(plan, nullptr);
                            HIP_CHECK(hipEventRecord(e0, nullptr));
                            for (int r = 0; r < reps && rc == 0; ++r) rc = hbs_grouped_run(plan, nullptr);
                            HIP_CHECK(hipEventRecord(e1, nullptr));
                            HIP_CHECK(hipEventSynchronize(e1));
                            if (rc) {
                                fprintf(stderr, "run failed: %s\n", hbs_strerror(rc));
                                return 2;
                            }
                            float ms = 0.0f;
                            HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                            const double per = ms / reps;
                            // time one work item holds a workgroup slot
                            const double ns_item = per * 1e6 * info.workgroups / (double)info.work_items;
                            if (per < best.ms) best = Probe{tile, split, per, ns_item};
                            hbs_grouped_plan_destroy(plan);
                        }
                    }
                    const double tflops = 2.0 * mc * nc * kc * count / (best.ms * 1e-3) / 1e12;
                    fprintf(out, "%s %d %d %d %d %d %.1f\n", pname, mc, nc, kc, best.tile,
                            best.split, best.ns_item);
                    printf("%s M%-5d N%-5d K%-5d -> tile %d split %d  %.3f ms  %.1f TFLOP/s\n",
                           pname, mc, nc, kc, best.tile, best.split, best.ms, tflops);
                    fflush(stdout);
                }
            }
        }
    }
    fclose(out);
    HIP_CHECK(hipFree(A));
    HIP_CHECK(hipFree(B));
    HIP_CHECK(hipFree(C));
    return 0;
}
