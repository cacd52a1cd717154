/* Exhaustive check of the orientation binning's division by 2pi
 * (div_two_pi in sift-wgpu_amd/csrc/kernels_detect.hip): for every f32 x in
 * [0, 36 * 2pi], y = x * RN(1/2pi) followed by one FMA correction must have the
 * floor of the IEEE quotient x / 2pi (the bin the reference computes). The
 * quotient itself may differ from the IEEE one in the last place; the counts of
 * such x, and of x whose floor the product alone would change, are printed.
 * Prints "DIV_TWO_PI_OK" and exits 0 when every floor agrees. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static float from_bits(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}

int main(void) {
    const float c = 6.28318548202514648438f; /* 2 * PI as f32, the kernels' kTwoPi */
    const float rc = 1.0f / c;
    const float hi = 36.0f * c;
    uint32_t last;
    memcpy(&last, &hi, 4);
    long long floor_bad = 0, quot_diff = 0, plain_bad = 0, n = 0;
#pragma omp parallel for reduction(+ : floor_bad, quot_diff, plain_bad, n) schedule(static, 1 << 20)
    for (long long i = 0; i <= (long long)last; i++) {
        const float x = from_bits((uint32_t)i);
        volatile float vx = x; /* keep the IEEE division a runtime division */
        const float ref = vx / c;
        const float y = vx * rc;
        const float q = fmaf(fmaf(-y, c, vx), rc, y);
        floor_bad += floorf(q) != floorf(ref);
        quot_diff += q != ref;
        plain_bad += floorf(y) != floorf(ref);
        n++;
    }
    printf("values %lld, floor differs %lld (product alone: %lld), quotient differs in the last place %lld\n",
           n, floor_bad, plain_bad, quot_diff);
    if (floor_bad || !plain_bad) return 1;
    printf("DIV_TWO_PI_OK\n");
    return 0;
}
