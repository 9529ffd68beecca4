// ssblas_example.cpp -- command-line driver with the argument grammar and
// output lines of the reference's example (example/main.cpp:189-267,
// print_perf :20-33, check_ans_batch.hpp), running ssblasGemmBatchedEx on
// the GPU through the C ABI.
//
//   ssblas_example INT_TYPE SCALE_TYPE AB_TYPE C_TYPE TRANSA TRANSB BATCH alpha M N K beta [lda ldb ldc]
//     INT_TYPE: int | longint     SCALE_TYPE: float | double | int
//     AB_TYPE:  float | double | char     C_TYPE: float | double | int
//
// Prints, like the reference example: the baseline start/end lines, the
// library call's status, "MAX ERROR <e>" (relative for fp, absolute for
// int) between the host baseline and the library, then one
// "M,N,K,time,[s],Gflops,[Gflops]" line for the baseline and one for the
// library call.  The baseline is a host triple loop (OpenMP over instances
// and columns) accumulating in the C type; the library call is timed
// around the synchronous entry point with the data already in device
// memory (A[]/B[]/C[] are device arrays of device pointers).
#include <hip/hip_runtime_api.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <cmath>
#include <random>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/ssblasBatch.h"

namespace {

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void print_perf(long M, long N, long K, long batch, double t)
{
    const double gflops = 2.0 * (double)M * (double)N * (double)K * (double)batch / 1e9 / t;
    printf("%ld,%ld,%ld,%f,[s],%f,[Gflops]\n", M, N, K, t, gflops);
}

[[noreturn]] void usage()
{
    printf("./a.out INT_TYPE SCALE_TYPE AB_TYPE C_TYPE TRANSA TRANSB BATCH alpha M N K beta\n");
    printf("./a.out INT_TYPE SCALE_TYPE AB_TYPE C_TYPE TRANSA TRANSB BATCH alpha M N K beta lda ldb ldc\n");
    printf("INT_TYPE: int longint\n");
    printf("SCALE_TYPE: float int\n");
    printf("AB_TYPE: float char\n");
    printf("C_TYPE: float int\n");
    exit(-1);
}

int type_of(const std::string& s, bool allow_char, const char* what)
{
    if (s == "float") return SSBLAS_R_32F;
    if (s == "double") return SSBLAS_R_64F;
    if (s == "int" && std::string(what) != "AB") return SSBLAS_R_32I;
    if (s == "char" && allow_char) return SSBLAS_R_8I;
    printf("UNKNOWN %s TYPE %s\n", what, s.c_str());
    exit(-1);
}

ssblasOperation_t trans_of(const std::string& s)
{
    if (s == "N") return SSBLAS_OP_N;
    if (s == "T") return SSBLAS_OP_T;
    if (s == "C") return SSBLAS_OP_C;
    printf("UNKNOWN TRANS %s\n", s.c_str());
    exit(-1);
}

#define HIP_OK(x)                                                                    \
    do {                                                                             \
        hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess) {                                                      \
            fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));           \
            exit(1);                                                                 \
        }                                                                            \
    } while (0)

// Host values of the reference example's distributions: fp U[0,1), int8
// operands U{1..32}, int32 C U{-64..64}; one engine per operand role.
template <class T>
void fill(std::vector<std::vector<T>>& M, long rows, long cols, long ld, unsigned seed)
{
    std::mt19937_64 eng(seed);
    for (auto& m : M)
        for (long j = 0; j < cols; ++j)
            for (long i = 0; i < rows; ++i) {
                if constexpr (std::is_same<T, float>::value || std::is_same<T, double>::value)
                    m[i + j * ld] = std::uniform_real_distribution<T>(0, 1)(eng);
                else if constexpr (std::is_same<T, signed char>::value)
                    m[i + j * ld] = (signed char)std::uniform_int_distribution<int>(1, 32)(eng);
                else
                    m[i + j * ld] = std::uniform_int_distribution<int>(-64, 64)(eng);
            }
}

template <class S, class AB, class CT>
int run(ssblasOperation_t ta, ssblasOperation_t tb, long batch, S alpha, long M, long N, long K, S beta, long lda,
        long ldb, long ldc, int abt, int ct)
{
    const long ca = ta == SSBLAS_OP_N ? K : M, ra = ta == SSBLAS_OP_N ? M : K;
    const long cb = tb == SSBLAS_OP_N ? N : K, rb = tb == SSBLAS_OP_N ? K : N;
    std::vector<std::vector<AB>> A(batch, std::vector<AB>(lda * ca)), B(batch, std::vector<AB>(ldb * cb));
    std::vector<std::vector<CT>> Cn(batch, std::vector<CT>(ldc * N));
    fill(A, ra, ca, lda, 10);
    fill(B, rb, cb, ldb, 11);
    fill(Cn, M, N, ldc, 12);
    std::vector<std::vector<CT>> Ct = Cn;

    printf("naive_gemm start\n");
    double t0 = now_s();
#pragma omp parallel for collapse(2) schedule(static)
    for (long b = 0; b < batch; ++b)
        for (long j = 0; j < N; ++j)
            for (long i = 0; i < M; ++i) {
                // int: exact sums, result modulo 2^32 like the library
                using Acc = typename std::conditional<std::is_same<CT, int>::value, long long, CT>::type;
                Acc acc = 0;
                for (long k = 0; k < K; ++k) {
                    const AB a = ta == SSBLAS_OP_N ? A[b][i + k * lda] : A[b][k + i * lda];
                    const AB x = tb == SSBLAS_OP_N ? B[b][k + j * ldb] : B[b][j + k * ldb];
                    acc += (Acc)a * (Acc)x;
                }
                CT& c = Cn[b][i + j * ldc];
                if constexpr (std::is_same<CT, int>::value)
                    c = (int)(uint32_t)((unsigned long long)alpha * (unsigned long long)acc +
                                        (unsigned long long)beta * (unsigned long long)(long long)c);
                else
                    c = (CT)(alpha * acc + beta * c);
            }
    const double t_naive = now_s() - t0;
    printf("naive_gemm end\n");

    // device copies and device pointer arrays
    std::vector<void*> dA(batch), dB(batch), dC(batch);
    for (long b = 0; b < batch; ++b) {
        HIP_OK(hipMalloc(&dA[b], sizeof(AB) * A[b].size()));
        HIP_OK(hipMalloc(&dB[b], sizeof(AB) * B[b].size()));
        HIP_OK(hipMalloc(&dC[b], sizeof(CT) * Ct[b].size()));
        HIP_OK(hipMemcpy(dA[b], A[b].data(), sizeof(AB) * A[b].size(), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(dB[b], B[b].data(), sizeof(AB) * B[b].size(), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(dC[b], Ct[b].data(), sizeof(CT) * Ct[b].size(), hipMemcpyHostToDevice));
    }
    void **pA, **pB, **pC;
    HIP_OK(hipMalloc(&pA, sizeof(void*) * batch));
    HIP_OK(hipMalloc(&pB, sizeof(void*) * batch));
    HIP_OK(hipMalloc(&pC, sizeof(void*) * batch));
    HIP_OK(hipMemcpy(pA, dA.data(), sizeof(void*) * batch, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(pB, dB.data(), sizeof(void*) * batch, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(pC, dC.data(), sizeof(void*) * batch, hipMemcpyHostToDevice));
    // one untimed call on a scratch copy of the first instance's C loads the
    // library's kernels, so the timed call measures the GEMM
    void *wC, **pW;
    HIP_OK(hipMalloc(&wC, sizeof(CT) * Ct[0].size()));
    HIP_OK(hipMemcpy(wC, Ct[0].data(), sizeof(CT) * Ct[0].size(), hipMemcpyHostToDevice));
    HIP_OK(hipMalloc(&pW, sizeof(void*)));
    HIP_OK(hipMemcpy(pW, &wC, sizeof(void*), hipMemcpyHostToDevice));
    (void)ssblasGemmBatchedEx(ta, tb, M, N, K, &alpha, (const void* const*)pA, (ssblasDataType_t)abt, lda,
                              (const void* const*)pB, (ssblasDataType_t)abt, ldb, &beta, (void* const*)pW,
                              (ssblasDataType_t)ct, ldc, 1, SSBLAS_COMPUTE_DEFAULT_TYPE, SSBLAS_COMPUTE_DEFAULT);
    HIP_OK(hipDeviceSynchronize());
    (void)hipFree(wC);
    (void)hipFree(pW);

    printf("ssblasGemmEx start\n");
    t0 = now_s();
    ssblasStatus_t st = ssblasGemmBatchedEx(ta, tb, M, N, K, &alpha, (const void* const*)pA, (ssblasDataType_t)abt,
                                            lda, (const void* const*)pB, (ssblasDataType_t)abt, ldb, &beta,
                                            (void* const*)pC, (ssblasDataType_t)ct, ldc, batch,
                                            SSBLAS_COMPUTE_DEFAULT_TYPE, SSBLAS_COMPUTE_DEFAULT);
    const double t_lib = now_s() - t0;
    printf("retval: %d\n", (int)st);
    ssblasShowError(st);
    for (long b = 0; b < batch; ++b)
        HIP_OK(hipMemcpy(Ct[b].data(), dC[b], sizeof(CT) * Ct[b].size(), hipMemcpyDeviceToHost));

    if constexpr (std::is_same<CT, int>::value) {
        long err = 0;
        for (long b = 0; b < batch; ++b)
            for (long j = 0; j < N; ++j)
                for (long i = 0; i < M; ++i) {
                    long d = std::labs((long)Cn[b][i + j * ldc] - (long)Ct[b][i + j * ldc]);
                    err = d > err ? d : err;
                }
        printf("MAX ERROR %d\n", (int)err);
    } else {
        CT err = 0;
        for (long b = 0; b < batch; ++b)
            for (long j = 0; j < N; ++j)
                for (long i = 0; i < M; ++i) {
                    const CT o = Cn[b][i + j * ldc];
                    CT d = o - Ct[b][i + j * ldc];
                    if (o != (CT)0) d /= o;
                    d = std::fabs(d);
                    err = d > err ? d : err;
                }
        printf("MAX ERROR %e\n", (double)err);
    }
    print_perf(M, N, K, batch, t_naive);
    print_perf(M, N, K, batch, t_lib);

    for (long b = 0; b < batch; ++b) {
        (void)hipFree(dA[b]);
        (void)hipFree(dB[b]);
        (void)hipFree(dC[b]);
    }
    (void)hipFree(pA);
    (void)hipFree(pB);
    (void)hipFree(pC);
    return st == SSBLAS_STATUS_SUCCESS ? 0 : 1;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 13 && argc != 16) usage();
    const std::string it = argv[1];
    if (it != "int" && it != "longint") {
        printf("UNKNOWN INT TYPE %s\n", it.c_str());
        return -1;
    }
    const int st = type_of(argv[2], true, "SCALE");
    const int abt = type_of(argv[3], true, "AB");
    const int ct = type_of(argv[4], false, "C");
    const ssblasOperation_t ta = trans_of(argv[5]), tb = trans_of(argv[6]);
    const long batch = atol(argv[7]);
    const char* alpha = argv[8];
    const long M = atol(argv[9]), N = atol(argv[10]), K = atol(argv[11]);
    const char* beta = argv[12];
    long lda, ldb, ldc;
    if (argc == 16) {
        lda = atol(argv[13]);
        ldb = atol(argv[14]);
        ldc = atol(argv[15]);
    } else {
        lda = ta == SSBLAS_OP_N ? M : K;
        ldb = tb == SSBLAS_OP_N ? K : N;
        ldc = M;
    }
    if (st == SSBLAS_R_32F && abt == SSBLAS_R_32F && ct == SSBLAS_R_32F)
        return run<float, float, float>(ta, tb, batch, (float)atof(alpha), M, N, K, (float)atof(beta), lda, ldb,
                                        ldc, abt, ct);
    if (st == SSBLAS_R_64F && abt == SSBLAS_R_64F && ct == SSBLAS_R_64F)
        return run<double, double, double>(ta, tb, batch, atof(alpha), M, N, K, atof(beta), lda, ldb, ldc, abt, ct);
    if (st == SSBLAS_R_32I && abt == SSBLAS_R_8I && ct == SSBLAS_R_32I)
        return run<int, signed char, int>(ta, tb, batch, atoi(alpha), M, N, K, atoi(beta), lda, ldb, ldc, abt, ct);
    printf("UNSUPPORTED TYPE COMBINATION\n");
    return -1;
}
