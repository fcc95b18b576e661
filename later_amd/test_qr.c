/*
 * test_qr -- command-line driver of the QR library, in plain C against
 * include/later.h, linked to liblater_amd.so.
 *
 * The reference's driver is `test_qr algo m n [-check]` (test/test_qr.cu:22-41):
 * algo 1 factors a uniform random m x n matrix with later_rgsqrf, prints the
 * time and GFLOPS = 2 n^2 (m - n/3) / (ms 1e6) (:80-86) and, with -check, the
 * orthogonality ||I - Q^T Q||_F / n and the backward error ||A - QR|| / ||A||
 * (:216-268).  This driver keeps that interface; only algo 1 (rgsqrf) is part
 * of this library.  Extra options:
 *   -form reference|fp32|mixed   options of later_rgsqrf_ex (default: the
 *                                reference's own entry point later_rgsqrf with
 *                                the reference's workspace sizes, :73-78)
 *   -normal                      N(0,1) entries instead of 2u - 1
 *   -seed S                      generator seed (default 0)
 *   -oc MIB                      out of core, as test/test_oc_qr.cu:69-75 drives
 *                                later_oc_qr_rec: A and R in pinned host memory,
 *                                at most MIB MiB of device memory (options of
 *                                -form, default reference); the time includes
 *                                the host-device copies
 * The last line is machine-readable:
 *   result algo=1 m=.. n=.. form=.. ms=.. gflops=.. orth=.. orth_over_n=.. backward=.. status=..
 * Exit status: 0 on success, 1 on a library or runtime error, 2 on bad usage.
 */
#include <hip/hip_runtime_api.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/later.h"

#define CHECK_HIP(x)                                                          \
    do {                                                                      \
        hipError_t e_ = (x);                                                  \
        if (e_ != hipSuccess) {                                               \
            fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_), \
                    __FILE__, __LINE__);                                      \
            return 1;                                                         \
        }                                                                     \
    } while (0)

static void usage(void)
{
    printf("Usage: test_qr algo m n [options]\n");
    printf("Options:\n\t-check: enable checking the orthogonality and backward error\n");
    printf("\t-form reference|fp32|mixed\n\t-normal\n\t-seed S\n\t-oc MIB\n");
}

int main(int argc, char **argv)
{
    if (argc < 4) {
        usage();
        return 2;
    }
    const int algo = atoi(argv[1]);
    const int m = atoi(argv[2]);
    const int n = atoi(argv[3]);
    int check = 0, normal = 0;
    long oc_mib = 0;
    unsigned seed = 0;
    const char *form = NULL;
    for (int i = 4; i < argc; ++i) {
        if (!strcmp(argv[i], "-check")) check = 1;
        else if (!strcmp(argv[i], "-normal")) normal = 1;
        else if (!strcmp(argv[i], "-seed") && i + 1 < argc) seed = (unsigned)strtoul(argv[++i], NULL, 10);
        else if (!strcmp(argv[i], "-form") && i + 1 < argc) form = argv[++i];
        else if (!strcmp(argv[i], "-oc") && i + 1 < argc) oc_mib = strtol(argv[++i], NULL, 10);
        else {
            usage();
            return 2;
        }
    }
    if (algo != 1) {
        printf("algo %d is not part of this library (only 1: rgsqrf)\n", algo);
        return 2;
    }
    if (m < 1 || n < 1 || m < n) {
        printf("bad shape %d x %d (need m >= n >= 1)\n", m, n);
        return 2;
    }
    later_opts o;
    if (form) {
        if (!strcmp(form, "reference")) later_opts_default(&o);
        else if (!strcmp(form, "fp32")) later_opts_fp32(&o);
        else if (!strcmp(form, "mixed")) later_opts_mixed(&o);
        else {
            usage();
            return 2;
        }
    }
    printf("%s\n", later_version());

    hipStream_t s;
    CHECK_HIP(hipStreamCreate(&s));
    laterCtxt ctxt = {(void *)s, NULL};
    float *A = NULL, *A0 = NULL, *R = NULL, *work = NULL;
    uint16_t *hwork = NULL;
    CHECK_HIP(hipMalloc((void **)&A, sizeof(float) * (size_t)m * n));
    CHECK_HIP(hipMalloc((void **)&R, sizeof(float) * (size_t)n * n));
    int rc = normal ? later_generate_normal(ctxt, m, n, A, m, seed, 0) : later_generate_uniform(ctxt, m, n, A, m, seed, 0);
    if (rc) {
        fprintf(stderr, "generator status %d\n", rc);
        return 1;
    }

    if (oc_mib > 0) {
        /* Out of core: the matrix goes to pinned host memory and comes back
         * factored; the device holds at most oc_mib MiB meanwhile. */
        if (!form) later_opts_default(&o);
        float *hA = NULL, *hR = NULL;
        CHECK_HIP(hipHostMalloc((void **)&hA, sizeof(float) * (size_t)m * n, 0));
        CHECK_HIP(hipHostMalloc((void **)&hR, sizeof(float) * (size_t)n * n, 0));
        CHECK_HIP(hipMemcpyAsync(hA, A, sizeof(float) * (size_t)m * n, hipMemcpyDeviceToHost, s));
        CHECK_HIP(hipStreamSynchronize(s));
        printf("Perform OC_QR_REC\nmatrix size %d*%d, device budget %ld MiB\n", m, n, oc_mib);
        hipEvent_t f0, f1;
        CHECK_HIP(hipEventCreate(&f0));
        CHECK_HIP(hipEventCreate(&f1));
        CHECK_HIP(hipEventRecord(f0, s));
        later_info info;
        const int st = later_oc_qr_rec(ctxt, m, n, hA, m, hR, n, (int64_t)oc_mib << 20, &o, &info);
        CHECK_HIP(hipEventRecord(f1, s));
        CHECK_HIP(hipEventSynchronize(f1));
        float ms = 0.0f;
        CHECK_HIP(hipEventElapsedTime(&ms, f0, f1));
        if (st) {
            fprintf(stderr, "later_oc_qr_rec status %d\n", st);
            return 1;
        }
        const double gflops = 2.0 * n * n * (m - n / 3.0) / (ms * 1e6);
        printf("OC_QR_REC takes %.3f ms, exec rate %.0f GFLOPS\n", ms, gflops);
        double orth = -1.0, bwd = -1.0;
        if (check) {
            CHECK_HIP(hipMemcpy(A, hA, sizeof(float) * (size_t)m * n, hipMemcpyHostToDevice));
            CHECK_HIP(hipMemcpy(R, hR, sizeof(float) * (size_t)n * n, hipMemcpyHostToDevice));
            CHECK_HIP(hipMalloc((void **)&A0, sizeof(float) * (size_t)m * n));
            rc = normal ? later_generate_normal(ctxt, m, n, A0, m, seed, 0)
                        : later_generate_uniform(ctxt, m, n, A0, m, seed, 0);
            if (!rc) rc = later_check_orth(ctxt, m, n, A, m, &orth);
            if (!rc) rc = later_check_backward(ctxt, m, n, A0, m, A, m, R, n, &bwd);
            if (rc) {
                fprintf(stderr, "check status %d\n", rc);
                return 1;
            }
            printf("orthogonality ||I - Q^T Q||_F / n = %.6e (raw %.6e)\n", orth / n, orth);
            printf("backward error ||A - QR||_F / ||A||_F = %.6e\n", bwd);
            CHECK_HIP(hipFree(A0));
        }
        printf("result algo=1 m=%d n=%d form=%s oc_mib=%ld ms=%.4f gflops=%.2f orth=%.6e orth_over_n=%.6e "
               "backward=%.6e status=%d\n",
               m, n, form ? form : "reference", oc_mib, ms, gflops, orth, orth >= 0 ? orth / n : -1.0, bwd,
               info.status);
        CHECK_HIP(hipHostFree(hA));
        CHECK_HIP(hipHostFree(hR));
        CHECK_HIP(hipFree(A));
        CHECK_HIP(hipFree(R));
        CHECK_HIP(hipEventDestroy(f0));
        CHECK_HIP(hipEventDestroy(f1));
        CHECK_HIP(hipStreamDestroy(s));
        return 0;
    }

    int64_t lwork = 0, lhwork = 0;
    if (form) {
        rc = later_rgsqrf_workspace(m, n, &o, &lwork, &lhwork);
        if (rc) {
            fprintf(stderr, "workspace query status %d\n", rc);
            return 1;
        }
    } else {
        /* the reference driver's sizes (test/test_qr.cu:73-76) */
        lwork = (int64_t)(m / 256) * 32 * n;
        lhwork = (int64_t)m * n;
    }
    CHECK_HIP(hipMalloc((void **)&work, sizeof(float) * (size_t)(lwork > 0 ? lwork : 1)));
    if (lhwork > 0) CHECK_HIP(hipMalloc((void **)&hwork, sizeof(uint16_t) * (size_t)lhwork));
    printf("Perform RGSQRF\nmatrix size %d*%d\n", m, n);

    hipEvent_t e0, e1;
    CHECK_HIP(hipEventCreate(&e0));
    CHECK_HIP(hipEventCreate(&e1));
    CHECK_HIP(hipStreamSynchronize(s));
    CHECK_HIP(hipEventRecord(e0, s));
    int status = 0;
    if (form) {
        later_info info;
        status = later_rgsqrf_ex(ctxt, m, n, A, m, R, n, work, lwork, hwork, lhwork, &o, &info);
    } else {
        later_rgsqrf(ctxt, m, n, A, m, R, n, work, (int)lwork, hwork, (int)lhwork);
    }
    CHECK_HIP(hipEventRecord(e1, s));
    CHECK_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (status) {
        fprintf(stderr, "later_rgsqrf_ex status %d\n", status);
        return 1;
    }
    int dev_status = 0;
    /* the status word is work[0] when lwork >= 64, else hwork's first word */
    rc = later_read_status(ctxt, lwork >= 64 ? work : (const float *)hwork, &dev_status);
    if (rc) {
        fprintf(stderr, "read_status status %d\n", rc);
        return 1;
    }
    const double gflops = 2.0 * n * n * (m - n / 3.0) / (ms * 1e6);
    printf("RGSQRF takes %.3f ms, exec rate %.0f GFLOPS\n", ms, gflops);

    double orth = -1.0, bwd = -1.0;
    if (check) {
        rc = later_check_orth(ctxt, m, n, A, m, &orth);
        if (rc) {
            fprintf(stderr, "check_orth status %d\n", rc);
            return 1;
        }
        printf("orthogonality ||I - Q^T Q||_F / n = %.6e (raw %.6e)\n", orth / n, orth);
        CHECK_HIP(hipMalloc((void **)&A0, sizeof(float) * (size_t)m * n));
        rc = normal ? later_generate_normal(ctxt, m, n, A0, m, seed, 0) : later_generate_uniform(ctxt, m, n, A0, m, seed, 0);
        if (!rc) rc = later_check_backward(ctxt, m, n, A0, m, A, m, R, n, &bwd);
        if (rc) {
            fprintf(stderr, "check_backward status %d\n", rc);
            return 1;
        }
        printf("backward error ||A - QR||_F / ||A||_F = %.6e\n", bwd);
        CHECK_HIP(hipFree(A0));
    }
    printf("result algo=1 m=%d n=%d form=%s ms=%.4f gflops=%.2f orth=%.6e orth_over_n=%.6e backward=%.6e "
           "status=%d\n",
           m, n, form ? form : "legacy", ms, gflops, orth, orth >= 0 ? orth / n : -1.0, bwd, dev_status);
    CHECK_HIP(hipFree(A));
    CHECK_HIP(hipFree(R));
    CHECK_HIP(hipFree(work));
    if (hwork) CHECK_HIP(hipFree(hwork));
    CHECK_HIP(hipEventDestroy(e0));
    CHECK_HIP(hipEventDestroy(e1));
    CHECK_HIP(hipStreamDestroy(s));
    return 0;
}
