// C ABI of the stage preprocessing: plm_hip_resample_geometry,
// plm_hip_volume_resample, plm_hip_stage_resample, plm_hip_hist_eqsp,
// plm_hip_hist_vopt.
// Declared in include/plm_hip_bspline.h.

#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "plm_hip_bspline.h"
#include "bspline_internal.h"
#include "stage_kernels.h"

using plmhip::ResampleAxis;

namespace {

// ROUND_INT / ROUND_PLM_LONG (src/sys/plm_math.h:49-54): the float widened
// to double, +0.5, truncated.
int64_t
round_long (float x)
{
    return x >= 0 ? (int64_t) ((double) x + 0.5) : (int64_t) (-(-(double) x + 0.5));
}

// One axis of volume_resample_float_li / _nn (src/base/volume_resample.cxx:
// 26-135): the output position advances by repeated float addition of the
// spacing from the origin; its input coordinate (pos - origin) / spacing
// is rounded for the inside test and, for the trilinear weights, clamped
// as li_clamp does with dmax = dim - 1 (src/base/interpolate.cxx:18-41,
// 88-103).
void
axis_table (float o_out, float s_out, int64_t n_out, float o_in, float s_in, int64_t n_in,
    bool nearest, ResampleAxis* t)
{
    float pos = o_out;
    const int64_t dmax = n_in - 1;
    for (int64_t i = 0; i < n_out; i++, pos += s_out) {
        const float kf = (pos - o_in) / s_in;
        // a non-finite coordinate is outside (the reference's cast of it is undefined)
        const int64_t r = std::isfinite (kf) ? round_long (kf) : -1;
        ResampleAxis e {0, 0.f, 0.f, r >= 0 && r < n_in};
        if (!e.inside) {
            // the voxel is 0 whatever the other axes say; index 0 stays in bounds
        } else if (nearest) {
            e.idx = (int) r;
        } else {
            int64_t f;
            float w2;
            if (kf < 0.f) {
                f = 0;
                w2 = 0.0f;
            } else if (kf >= dmax) {
                f = dmax - 1;
                w2 = 1.0f;
            } else {
                f = (int64_t) std::floor (kf);
                w2 = kf - f;
            }
            e.idx = (int) f;
            e.w2 = w2;
            e.w1 = 1.0f - w2;
        }
        t[i] = e;
    }
}

int
check_out_dim (const int64_t dim[3])
{
    for (int d = 0; d < 3; d++)
        if (dim[d] < 1) return plmhip::abi_fail (PLM_HIP_ERR_GEOMETRY, "output dim < 1");
    if (dim[1] > 65535 || dim[2] > 65535)
        return plmhip::abi_fail (PLM_HIP_ERR_GEOMETRY, "output dim[1], dim[2] must be <= 65535");
    if (dim[0] * dim[1] * dim[2] >= ((int64_t) 1 << 31))
        return plmhip::abi_fail (PLM_HIP_ERR_GEOMETRY, "output has too many voxels");
    return PLM_HIP_OK;
}

int
select_device (int gpuid)
{
    int ndev = 0;
    hipError_t e = hipGetDeviceCount (&ndev);
    if (e != hipSuccess || ndev < 1) return plmhip::abi_fail (PLM_HIP_ERR_NO_DEVICE, "no HIP device");
    if (gpuid < 0 || gpuid >= ndev) return plmhip::abi_fail (PLM_HIP_ERR_ARG, "gpuid out of range");
    e = hipSetDevice (gpuid);
    if (e != hipSuccess) return plmhip::abi_hip_fail (e, "hipSetDevice");
    return PLM_HIP_OK;
}

// Device buffers of one call, freed on every way out.
struct Bufs {
    std::vector<void*> p;
    ~Bufs ()
    {
        for (void* b : p) (void) hipFree (b);
    }
    template <typename T>
    int get (T** out, size_t count, const char* what)
    {
        hipError_t e = hipMalloc ((void**) out, sizeof (T) * (count ? count : 1));
        if (e != hipSuccess) return plmhip::abi_hip_fail (e, what);
        p.push_back (*out);
        return PLM_HIP_OK;
    }
};

// The input on the device (uploaded into bufs when it is on the host).
int
device_input (const plm_img_desc* in, Bufs& bufs, const float** src)
{
    const int64_t n = in->dim[0] * in->dim[1] * in->dim[2];
    if (in->location == PLM_MEM_DEVICE) {
        *src = (const float*) in->data;
        return PLM_HIP_OK;
    }
    float* d = nullptr;
    int rc = bufs.get (&d, (size_t) n, "hipMalloc(resample input)");
    if (rc) return rc;
    hipError_t e = hipMemcpy (d, in->data, sizeof (float) * n, hipMemcpyHostToDevice);
    if (e != hipSuccess) return plmhip::abi_hip_fail (e, "hipMemcpy(resample input)");
    *src = d;
    return PLM_HIP_OK;
}

// Minimum and maximum of a device image (the scan of
// bspline_initialize_mi_hist_eqsp / _vopt; a NaN fails both comparisons,
// as there).
int
image_range (const float* src, const plm_img_desc* vol, Bufs& bufs, float* mn, float* mx)
{
    float* d_part = nullptr;
    int rc = bufs.get (&d_part, 2 * (size_t) plmhip::MINMAX_BLOCKS, "hipMalloc(min/max partials)");
    if (rc) return rc;
    const int64_t n = vol->dim[0] * vol->dim[1] * vol->dim[2];
    int nb = 0;
    hipError_t e = plmhip::launch_minmax (src, n, d_part, &nb, nullptr);
    std::vector<float> part (2 * (size_t) nb);
    if (e == hipSuccess) e = hipMemcpy (part.data (), d_part, sizeof (float) * part.size (), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return plmhip::abi_hip_fail (e, "min/max scan");
    float a = part[0], b = part[1];
    for (int k = 1; k < nb; k++) {
        a = part[2 * k] < a ? part[2 * k] : a;
        b = part[2 * k + 1] > b ? part[2 * k + 1] : b;
    }
    *mn = a;
    *mx = b;
    return PLM_HIP_OK;
}

} // namespace

extern "C" {

int
plm_hip_resample_geometry (int kind, const plm_img_desc* in, const float rate[3], int64_t dim[3],
    float origin[3], float spacing[3], int* identity)
{
    int rc = plmhip::abi_check_img (in, "in");
    if (rc) return rc;
    if (!rate || !dim || !origin || !spacing || !identity)
        return plmhip::abi_fail (PLM_HIP_ERR_ARG, "NULL argument");
    *identity = 0;
    if (kind == PLM_RESAMPLE_VOXEL_RATE || kind == PLM_RESAMPLE_VOXEL_LEGACY) {
        if (rate[0] == 1.0f && rate[1] == 1.0f && rate[2] == 1.0f) {
            for (int d = 0; d < 3; d++) {
                dim[d] = in->dim[d];
                origin[d] = in->origin[d];
                spacing[d] = in->spacing[d];
            }
            *identity = 1;
            return PLM_HIP_OK;
        }
    } else if (kind != PLM_RESAMPLE_MM && kind != PLM_RESAMPLE_PCT) {
        return plmhip::abi_fail (PLM_HIP_ERR_ARG, "unknown resample kind");
    }
    for (int d = 0; d < 3; d++) {
        switch (kind) {
        case PLM_RESAMPLE_VOXEL_RATE: {
            // volume_subsample_vox (src/base/volume_resample.cxx:298-315)
            const int64_t int_rate = round_long (rate[d]);
            if (int_rate < 1) return plmhip::abi_fail (PLM_HIP_ERR_ARG, "voxel rate rounds to 0");
            dim[d] = (in->dim[d] + int_rate - 1) / int_rate;
            origin[d] = in->origin[d];
            spacing[d] = in->spacing[d] * int_rate;
            break;
        }
        case PLM_RESAMPLE_VOXEL_LEGACY: {
            // volume_subsample_vox_legacy (:338-357)
            if ((int64_t) rate[d] < 1) return plmhip::abi_fail (PLM_HIP_ERR_ARG, "voxel rate truncates to 0");
            const float in_size = in->dim[d] * in->spacing[d];
            dim[d] = in->dim[d] / (int64_t) rate[d];
            if (dim[d] < 1) dim[d] = 1;
            spacing[d] = in_size / dim[d];
            origin[d] = (float) (in->origin[d] - 0.5 * in->spacing[d] + 0.5 * spacing[d]);
            break;
        }
        case PLM_RESAMPLE_MM: {
            // volume_resample_spacing (:382-393)
            if (!(rate[d] > 0.f)) return plmhip::abi_fail (PLM_HIP_ERR_ARG, "spacing must be positive");
            const float in_size = (in->dim[d] - 1) * in->spacing[d];
            dim[d] = 1 + round_long (in_size / rate[d]);
            origin[d] = in->origin[d];
            spacing[d] = rate[d];
            break;
        }
        default: {
            // volume_resample_percent (:396-413)
            const float in_size = (in->dim[d] - 1) * in->spacing[d];
            dim[d] = 1 + round_long (rate[d] * (in->dim[d] - 1));
            spacing[d] = dim[d] == 1 ? in_size : in_size / (dim[d] - 1);
            origin[d] = in->origin[d];
            break;
        }
        }
    }
    return PLM_HIP_OK;
}

int
plm_hip_volume_resample (const plm_img_desc* in, const int64_t dim[3], const float origin[3],
    const float spacing[3], int nearest, float* out, int out_location, int gpuid)
{
    int rc = plmhip::abi_check_img (in, "in");
    if (rc) return rc;
    if (!dim || !origin || !spacing || !out) return plmhip::abi_fail (PLM_HIP_ERR_ARG, "NULL argument");
    if (in->storage != PLM_STORE_F32) return plmhip::abi_fail (PLM_HIP_ERR_ARG, "in must be float32");
    if (out_location != PLM_MEM_HOST && out_location != PLM_MEM_DEVICE)
        return plmhip::abi_fail (PLM_HIP_ERR_ARG, "bad out_location");
    if ((rc = check_out_dim (dim)) || (rc = select_device (gpuid))) return rc;
    std::vector<ResampleAxis> tab ((size_t) (dim[0] + dim[1] + dim[2]));
    for (int d = 0, off = 0; d < 3; off += (int) dim[d], d++)
        axis_table (origin[d], spacing[d], dim[d], in->origin[d], in->spacing[d], in->dim[d], nearest != 0,
            tab.data () + off);
    Bufs bufs;
    const float* src = nullptr;
    ResampleAxis* d_tab = nullptr;
    float* d_out = nullptr;
    const int64_t nout = dim[0] * dim[1] * dim[2];
    if ((rc = device_input (in, bufs, &src))
        || (rc = bufs.get (&d_tab, tab.size (), "hipMalloc(resample tables)")))
        return rc;
    if (out_location == PLM_MEM_HOST && (rc = bufs.get (&d_out, (size_t) nout, "hipMalloc(resample output)")))
        return rc;
    hipError_t e = hipMemcpy (d_tab, tab.data (), sizeof (ResampleAxis) * tab.size (), hipMemcpyHostToDevice);
    if (e != hipSuccess) return plmhip::abi_hip_fail (e, "hipMemcpy(resample tables)");
    const int id[3] = {(int) in->dim[0], (int) in->dim[1], (int) in->dim[2]};
    const int od[3] = {(int) dim[0], (int) dim[1], (int) dim[2]};
    float* dst = d_out ? d_out : out;
    e = plmhip::launch_resample (src, id, dst, od, d_tab, d_tab + od[0], d_tab + od[0] + od[1], nearest != 0,
        nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize ();
    if (e == hipSuccess && d_out) e = hipMemcpy (out, d_out, sizeof (float) * nout, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return plmhip::abi_hip_fail (e, "volume resample");
    return PLM_HIP_OK;
}

int
plm_hip_stage_resample (int kind, const float rate[3], const plm_img_desc* in, int nearest, float* out,
    int out_location, int gpuid)
{
    int64_t dim[3];
    float origin[3], spacing[3];
    int identity = 0;
    int rc = plm_hip_resample_geometry (kind, in, rate, dim, origin, spacing, &identity);
    if (rc) return rc;
    if (!out) return plmhip::abi_fail (PLM_HIP_ERR_ARG, "out is NULL");
    if (!identity) return plm_hip_volume_resample (in, dim, origin, spacing, nearest, out, out_location, gpuid);
    // rate 1, 1, 1: the reference clones the volume
    if ((rc = select_device (gpuid))) return rc;
    const hipMemcpyKind k = in->location == PLM_MEM_HOST
        ? (out_location == PLM_MEM_HOST ? hipMemcpyHostToHost : hipMemcpyHostToDevice)
        : (out_location == PLM_MEM_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
    hipError_t e = hipMemcpy (out, in->data, sizeof (float) * dim[0] * dim[1] * dim[2], k);
    if (e != hipSuccess) return plmhip::abi_hip_fail (e, "hipMemcpy(stage resample clone)");
    return PLM_HIP_OK;
}

int
plm_hip_hist_eqsp (const plm_img_desc* vol, int64_t bins, float* offset, float* delta, int gpuid)
{
    int rc = plmhip::abi_check_img (vol, "vol");
    if (rc) return rc;
    if (!offset || !delta) return plmhip::abi_fail (PLM_HIP_ERR_ARG, "NULL argument");
    if (bins < 2) return plmhip::abi_fail (PLM_HIP_ERR_ARG, "bins must be >= 2");
    if (vol->storage != PLM_STORE_F32) return plmhip::abi_fail (PLM_HIP_ERR_ARG, "vol must be float32");
    if ((rc = select_device (gpuid))) return rc;
    Bufs bufs;
    const float* src = nullptr;
    float mn = 0.f, mx = 0.f;
    if ((rc = device_input (vol, bufs, &src)) || (rc = image_range (src, vol, bufs, &mn, &mx))) return rc;
    // bspline_initialize_mi_hist_eqsp (src/register/joint_histogram.cxx:
    // 131-133): the top and bottom bins are half full
    *delta = (mx - mn) / (float) (bins - 1);
    *offset = (float) (mn - 0.5 * *delta);
    return PLM_HIP_OK;
}

// bspline_initialize_mi_hist_vopt (src/register/joint_histogram.cxx:
// 195-324).  The key sums are the one place where the reference's order of
// additions could matter (it adds v - offset key by key in voxel order, in
// double); it does not: every term of key k is a float in
// [k delta, (k + 1) delta) -- key 0 starts at delta / 2, the image minimum
// -- so the terms of one key lie within a factor of two of each other and
// are multiples of the smallest one's ulp, 2^(e - 23); a sum of fewer than
// 2^28 of them stays below 2^(e + 30) and is itself such a multiple,
// exact in double's 53 bits.  Every order then gives the exact sum, and the
// device adds them with atomics.
int
plm_hip_hist_vopt (const plm_img_desc* vol, int64_t bins, float* offset, float* delta,
    int32_t* key_lut, int gpuid)
{
    constexpr int K = plmhip::VOPT_KEYS;
    int rc = plmhip::abi_check_img (vol, "vol");
    if (rc) return rc;
    if (!offset || !delta || !key_lut) return plmhip::abi_fail (PLM_HIP_ERR_ARG, "NULL argument");
    if (bins < 2 || bins > K)
        return plmhip::abi_fail (PLM_HIP_ERR_ARG, "bins must be in [2, PLM_HIP_VOPT_KEYS]");
    if (vol->storage != PLM_STORE_F32) return plmhip::abi_fail (PLM_HIP_ERR_ARG, "vol must be float32");
    const int64_t n = vol->dim[0] * vol->dim[1] * vol->dim[2];
    if (n >= ((int64_t) 1 << 28))
        return plmhip::abi_fail (PLM_HIP_ERR_UNSUPPORTED, "image too large for exact key sums (2^28 voxels)");
    if ((rc = select_device (gpuid))) return rc;
    Bufs bufs;
    const float* src = nullptr;
    float mn = 0.f, mx = 0.f;
    if ((rc = device_input (vol, bufs, &src)) || (rc = image_range (src, vol, bufs, &mn, &mx))) return rc;
    if (!(mx > mn)) return plmhip::abi_fail (PLM_HIP_ERR_ARG, "constant image: no histogram range");
    // the equal-spacing grid of K keys (src/register/joint_histogram.cxx:242-244)
    const float dl = (mx - mn) / (float) (K - 1);
    const float off = (float) (mn - 0.5 * dl);
    // one allocation: key counts, key sums (then the three prefix tables),
    // the programme's last two rows, the splits of every row
    char* blob = nullptr;
    if ((rc = bufs.get (&blob, sizeof (double) * 6 * K + sizeof (int) * (size_t) bins * K, "hipMalloc(VOPT tables)")))
        return rc;
    unsigned long long* d_cnt = reinterpret_cast<unsigned long long*> (blob);
    double* d_tab = reinterpret_cast<double*> (blob) + K;
    double* d_err = d_tab + 3 * K;
    int* d_split = reinterpret_cast<int*> (d_err + 2 * K);
    std::vector<unsigned long long> cnt1 (K);
    std::vector<double> sum (K);
    // counts and sums are adjacent: one clear, one read-back
    std::vector<unsigned long long> ks (2 * (size_t) K);
    hipError_t e = hipMemsetAsync (blob, 0, sizeof (double) * 2 * K, nullptr);
    if (e == hipSuccess) e = plmhip::launch_vopt_keys (src, n, off, dl, d_cnt, d_tab, nullptr);
    if (e == hipSuccess) e = hipMemcpy (ks.data (), blob, sizeof (double) * 2 * K, hipMemcpyDeviceToHost);
    if (e == hipSuccess) {
        std::memcpy (cnt1.data (), ks.data (), sizeof (unsigned long long) * K);
        std::memcpy (sum.data (), ks.data () + K, sizeof (double) * K);
    }
    if (e != hipSuccess) return plmhip::abi_hip_fail (e, "VOPT key sums");
    // key averages (an empty key repeats the previous one's) and their
    // running sums, in key order (:253-272)
    std::vector<double> avg (K), tab (3 * K);
    for (int i = 0; i < K; i++) {
        if (cnt1[i] > 0) avg[i] = sum[i] / (double) cnt1[i];
        else avg[i] = i > 0 ? avg[i - 1] : 0.0;
    }
    double* ps = tab.data ();
    double* pssq = ps + K;
    double* pcnt = pssq + K;
    ps[0] = avg[0];
    pssq[0] = avg[0] * avg[0];
    pcnt[0] = (double) cnt1[0];
    for (int i = 1; i < K; i++) {
        ps[i] = ps[i - 1] + avg[i];
        pssq[i] = pssq[i - 1] + avg[i] * avg[i];
        pcnt[i] = pcnt[i - 1] + (double) cnt1[i];
    }
    std::vector<int> split ((size_t) bins * K);
    e = hipMemcpy (d_tab, tab.data (), sizeof (double) * 3 * K, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = plmhip::launch_vopt_dp (d_tab, d_tab + K, d_tab + 2 * K, (int) bins, d_err, d_split, nullptr);
    if (e == hipSuccess) e = hipMemcpy (split.data (), d_split, sizeof (int) * split.size (), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return plmhip::abi_hip_fail (e, "VOPT partition");
    // the key table, walking the best splits back from the last key (:304-321)
    for (int i = 0; i < K; i++) key_lut[i] = 0;
    int hi = K - 1;
    for (int j = (int) bins - 1; j >= 0; j--) {
        const int nx = j > 0 ? split[(size_t) j * K + hi] : 0;
        const int lo = j == 0 ? 0 : nx + 1;
        for (int i = lo; i <= hi; i++) key_lut[i] = j;
        hi = nx;
    }
    *offset = off;
    *delta = dl;
    return PLM_HIP_OK;
}

} // extern "C"
