// Records v_mfma_scale_f32_16x16x128_f8f6f4 (e4m3fn operands, FP32 accumulator) on
// given operands: the golden vectors of the oracle's restatement of the FP8 candidate
// pass (oracle/ssg_oracle.c or_fp8_mfma).  Driven by make_fp8_mfma.py:
//   fp8_mfma_cases IN OUT
// IN: int32 ncases, then per case int32 scale_a, int32 scale_b (E8M0), uint8 A[16][128],
// uint8 B[128][16], float C[16][16].  OUT: per case float D[16][16], D = A B 2^s + C with
// s = scale_a + scale_b - 254.  One wave per case: lane l holds A row l % 16 and B column
// l % 16 at k in [32 (l / 16), +32), and D rows 4 (l / 16) + i, column l % 16.
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <vector>

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef int intx8 __attribute__((ext_vector_type(8)));

__global__ void mfma_case(const uint8_t* A, const uint8_t* B, const float* C, int sa, int sb, float* D)
{
    const int l = threadIdx.x;
    uint8_t ta[32], tb[32];
    for (int e = 0; e < 32; e++) {
        ta[e] = A[(l % 16) * 128 + 32 * (l / 16) + e];
        tb[e] = B[(32 * (l / 16) + e) * 16 + (l % 16)];
    }
    intx8 a, b;
    __builtin_memcpy(&a, ta, 32);
    __builtin_memcpy(&b, tb, 32);
    floatx4 c;
    for (int i = 0; i < 4; i++) c[i] = C[(4 * (l / 16) + i) * 16 + l % 16];
    c = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a, b, c, 0, 0, 0, sa, 0, sb);
    for (int i = 0; i < 4; i++) D[(4 * (l / 16) + i) * 16 + l % 16] = c[i];
}

int main(int argc, char** argv)
{
    if (argc != 3) {
        fprintf(stderr, "usage: %s IN OUT\n", argv[0]);
        return 2;
    }
    FILE* f = fopen(argv[1], "rb");
    if (!f) return 2;
    int32_t nc = 0;
    if (fread(&nc, 4, 1, f) != 1 || nc <= 0 || nc > 1024) return 2;
    std::vector<float> out((size_t)nc * 256);
    uint8_t *dA, *dB;
    float *dC, *dD;
    if (hipMalloc(&dA, 2048) || hipMalloc(&dB, 2048) || hipMalloc(&dC, 1024) || hipMalloc(&dD, 1024)) return 1;
    for (int c = 0; c < nc; c++) {
        int32_t s[2];
        uint8_t hA[2048], hB[2048];
        float hC[256];
        if (fread(s, 4, 2, f) != 2 || fread(hA, 1, 2048, f) != 2048 || fread(hB, 1, 2048, f) != 2048 ||
            fread(hC, 4, 256, f) != 256)
            return 2;
        if (s[0] < 0 || s[0] > 254 || s[1] < 0 || s[1] > 254) return 2;
        if (hipMemcpy(dA, hA, 2048, hipMemcpyHostToDevice) || hipMemcpy(dB, hB, 2048, hipMemcpyHostToDevice) ||
            hipMemcpy(dC, hC, 1024, hipMemcpyHostToDevice))
            return 1;
        hipLaunchKernelGGL(mfma_case, dim3(1), dim3(64), 0, 0, dA, dB, dC, s[0], s[1], dD);
        if (hipDeviceSynchronize() || hipMemcpy(out.data() + (size_t)c * 256, dD, 1024, hipMemcpyDeviceToHost))
            return 1;
    }
    fclose(f);
    FILE* g = fopen(argv[2], "wb");
    if (!g || fwrite(out.data(), 4, out.size(), g) != out.size()) return 2;
    fclose(g);
    printf("%d cases recorded\n", nc);
    return 0;
}
