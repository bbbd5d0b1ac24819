/* Stage subsampling: CPU restatement (TEST INFRASTRUCTURE).
 *
 * registration_resample_volume (src/register/registration_resample.cxx:
 * 13-58) picks an output geometry by the stage's resample type --
 * volume_subsample_vox / _legacy, volume_resample_spacing / _percent
 * (src/base/volume_resample.cxx:298-413) -- and resamples the float volume
 * onto it, trilinearly (volume_resample_float_li, :71-135) or, for the
 * masks and stiffness maps, by the nearest voxel (volume_resample_float_nn,
 * :26-67).  Only the tests and the CPU leg of bench.py load this.
 *
 * Float expressions keep the reference's operation order (the file is
 * compiled with -ffp-contract=off): positions advance by repeated addition
 * of the output spacing, ROUND_INT rounds in double
 * (src/sys/plm_math.h:49), li_clamp takes dim - 1 as its bound
 * (src/base/interpolate.cxx:18-41, 88-103) and LI_VALUE sums the eight
 * weighted corners left to right (src/base/interpolate_macros.h:23-38).
 */
#include <math.h>
#include <stdint.h>

#include "plm_oracle.h"

/* ROUND_INT / ROUND_PLM_LONG: the float widened to double, +0.5, truncated */
static int64_t
round_long (float x)
{
    return x >= 0 ? (int64_t) ((double) x + 0.5) : (int64_t) (-(-(double) x + 0.5));
}

int
orc_resample_geometry (int kind, const orc_vol_hdr* in, const float rate[3], int64_t dim[3],
    float origin[3], float spacing[3])
{
    if (kind == ORC_RESAMPLE_VOXEL_RATE || kind == ORC_RESAMPLE_VOXEL_LEGACY) {
        if (rate[0] == 1.0f && rate[1] == 1.0f && rate[2] == 1.0f) {
            for (int d = 0; d < 3; d++) {
                dim[d] = in->dim[d];
                origin[d] = in->origin[d];
                spacing[d] = in->spacing[d];
            }
            return 1;       /* the volume is cloned, not resampled */
        }
    }
    for (int d = 0; d < 3; d++) {
        /* a rate that truncates or rounds to 0 divides by zero in the
           reference; refused here */
        if ((kind == ORC_RESAMPLE_VOXEL_RATE && round_long (rate[d]) < 1)
            || (kind == ORC_RESAMPLE_VOXEL_LEGACY && (int64_t) rate[d] < 1)
            || (kind == ORC_RESAMPLE_MM && !(rate[d] > 0.f)))
            return -1;
        switch (kind) {
        case ORC_RESAMPLE_VOXEL_RATE: {
            const int64_t int_rate = round_long (rate[d]);
            dim[d] = (in->dim[d] + int_rate - 1) / int_rate;
            origin[d] = in->origin[d];
            spacing[d] = in->spacing[d] * int_rate;
            break;
        }
        case ORC_RESAMPLE_VOXEL_LEGACY: {
            const float in_size = in->dim[d] * in->spacing[d];
            dim[d] = in->dim[d] / (int64_t) rate[d];
            if (dim[d] < 1) dim[d] = 1;
            spacing[d] = in_size / dim[d];
            origin[d] = (float) (in->origin[d] - 0.5 * in->spacing[d] + 0.5 * spacing[d]);
            break;
        }
        case ORC_RESAMPLE_MM: {
            const float in_size = (in->dim[d] - 1) * in->spacing[d];
            dim[d] = 1 + round_long (in_size / rate[d]);
            origin[d] = in->origin[d];
            spacing[d] = rate[d];
            break;
        }
        case ORC_RESAMPLE_PCT: {
            const float in_size = (in->dim[d] - 1) * in->spacing[d];
            dim[d] = 1 + round_long (rate[d] * (in->dim[d] - 1));
            spacing[d] = dim[d] == 1 ? in_size : in_size / (dim[d] - 1);
            origin[d] = in->origin[d];
            break;
        }
        default:
            return -1;
        }
    }
    return 0;
}

/* li_clamp with dmax = dim - 1 */
static void
li_clamp_d (float ma, int64_t dmax, int64_t* maf, float* fa1, float* fa2)
{
    if (ma < 0.f) {
        *maf = 0;
        *fa2 = 0.0f;
    } else if (ma >= dmax) {
        *maf = dmax - 1;
        *fa2 = 1.0f;
    } else {
        *maf = (int64_t) floor (ma);
        *fa2 = ma - *maf;
    }
    *fa1 = 1.0f - *fa2;
}

void
orc_volume_resample (const orc_vol_hdr* in, const float* img, const int64_t dim[3],
    const float origin[3], const float spacing[3], int nearest, float* out)
{
    const int64_t d0 = in->dim[0], d01 = in->dim[0] * in->dim[1];
    int64_t v = 0;
    float z = origin[2];
    for (int64_t k = 0; k < dim[2]; k++, z += spacing[2]) {
        const float kz = (z - in->origin[2]) / in->spacing[2];
        const int64_t rz = round_long (kz);
        float y = origin[1];
        for (int64_t j = 0; j < dim[1]; j++, y += spacing[1]) {
            const float ky = (y - in->origin[1]) / in->spacing[1];
            const int64_t ry = round_long (ky);
            float x = origin[0];
            for (int64_t i = 0; i < dim[0]; i++, x += spacing[0], v++) {
                const float kx = (x - in->origin[0]) / in->spacing[0];
                const int64_t rx = round_long (kx);
                const int inside = rx >= 0 && rx < in->dim[0] && ry >= 0 && ry < in->dim[1]
                    && rz >= 0 && rz < in->dim[2];
                if (!inside) {
                    out[v] = 0.0f;
                    continue;
                }
                if (nearest) {
                    out[v] = img[rx + d0 * ry + d01 * rz];
                    continue;
                }
                int64_t fx, fy, fz;
                float x1, x2, y1, y2, z1, z2;
                li_clamp_d (kx, in->dim[0] - 1, &fx, &x1, &x2);
                li_clamp_d (ky, in->dim[1] - 1, &fy, &y1, &y2);
                li_clamp_d (kz, in->dim[2] - 1, &fz, &z1, &z2);
                const int64_t f = fx + d0 * (fy + in->dim[1] * fz);
                const float a = x1 * y1 * z1 * img[f];
                const float b = x2 * y1 * z1 * img[f + 1];
                const float c = x1 * y2 * z1 * img[f + d0];
                const float e = x2 * y2 * z1 * img[f + d0 + 1];
                const float g = x1 * y1 * z2 * img[f + d01];
                const float h = x2 * y1 * z2 * img[f + d01 + 1];
                const float p = x1 * y2 * z2 * img[f + d01 + d0];
                const float q = x2 * y2 * z2 * img[f + d01 + d0 + 1];
                out[v] = a + b + c + e + g + h + p + q;
            }
        }
    }
}
