// Upstream producers of the transport's inputs, on the device: the vertical
// water flux from continuity (ModuleHydrodynamic::WaterFlux_Z_VarGrid /
// _FixGrid), the cell volumes (ModuleGeometry::ComputeVolumes) and the face
// areas (ModuleGeometry::ComputeAreas).  Each entry point works on caller
// arrays over Size, in place on the outputs; with device_pointers the arrays
// stay on the device, so a caller can chain geometry, continuity and the
// batched transport without a host round trip.
#include <hip/hip_runtime.h>

#include <vector>

#include "mgpu_internal.h"

namespace {

// One thread per WorkSize column; the levels in order, each W face from the
// one below (HYD:43286-43302, 43325-43339).
__global__ void waterflux_z_kernel(Dims d, const double *__restrict__ vnew,
                                   const double *__restrict__ vold,
                                   const double *__restrict__ qx, const double *__restrict__ qy,
                                   double *__restrict__ qz, const int32_t *__restrict__ cfu,
                                   const int32_t *__restrict__ cfv,
                                   const int32_t *__restrict__ wp, double dt, int vargrid) {
    const int i = d.wILB + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = d.wJLB + blockIdx.y;
    if (i > d.wIUB) return;
    for (int k = d.wKLB; k <= d.wKUB; ++k) {
        const long x = d.idx(i, j, k);
        if (wp[x] != 1) continue;
        const long xe = x + d.sj, xn = x + 1;
        const double west = qx[x] * (double)cfu[x];
        const double east = qx[xe] * (double)cfu[xe];
        const double south = qy[x] * (double)cfv[x];
        const double north = qy[xn] * (double)cfv[xn];
        double v = qz[x] + west - east + south - north;
        if (vargrid) v = v - (vnew[x] - vold[x]) / dt;
        qz[x + d.sk] = v;
    }
}

// GEO:4137-4148: VolumeZ = DWZ * DUX * DVY over the WorkSize
__global__ void volume_z_kernel(Dims d, const double *__restrict__ dwz,
                                const double *__restrict__ dux, const double *__restrict__ dvy,
                                double *__restrict__ vz) {
    const int i = d.wILB + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = d.wJLB + blockIdx.y, k = d.wKLB + blockIdx.z;
    if (i > d.wIUB) return;
    const long x = d.idx(i, j, k);
    vz[x] = dwz[x] * dux[d.idx2(i, j)] * dvy[d.idx2(i, j)];
}

// GEO:4151-4212: the U, V (and W) volumes, means of the two Z cells beside
// each face
__global__ void volume_faces_kernel(Dims d, const double *__restrict__ vz,
                                    double *__restrict__ vu, double *__restrict__ vv,
                                    double *__restrict__ vw) {
    const int i = d.wILB + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = d.wJLB + blockIdx.y, k = d.wKLB + blockIdx.z;
    if (i > d.wIUB) return;
    const long x = d.idx(i, j, k);
    if (j > d.wJLB) vu[x] = (vz[x - d.sj] + vz[x]) / 2.0;
    if (i > d.wILB) vv[x] = (vz[x - 1] + vz[x]) / 2.0;
    if (vw && k > d.wKLB) vw[x] = (vz[x - d.sk] + vz[x]) / 2.0;
}

// GEO:4265-4305: AreaU = DUZ * DYY, AreaV = DVZ * DXX, then with
// impermeability the factors of the water level
__global__ void areas_kernel(Dims d, const double *__restrict__ duz,
                             const double *__restrict__ dvz, const double *__restrict__ dxx,
                             const double *__restrict__ dyy, double *__restrict__ au,
                             double *__restrict__ av, const double *__restrict__ szz,
                             const double *__restrict__ cu, const double *__restrict__ cxu,
                             const double *__restrict__ cv, const double *__restrict__ cxv) {
    const int i = d.wILB + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = d.wJLB + blockIdx.y, k = d.wKLB + blockIdx.z;
    if (i > d.wIUB) return;
    const long x = d.idx(i, j, k);
    double a = duz[x] * dyy[d.idx2(i, j)];
    double b = dvz[x] * dxx[d.idx2(i, j)];
    if (szz) {
        const int kc = k - d.KLB;
        const long s0 = d.idx(i, j, d.wKUB);
        const double su = -szz[s0 + d.sj], sv = -szz[s0 + 1], s = -szz[s0];
        double wl = s > su ? s : su;
        a = a * (cu[kc] + wl * cxu[kc]);
        wl = s > sv ? s : sv;
        b = b * (cv[kc] + wl * cxv[kc]);
    }
    au[x] = a;
    av[x] = b;
}

// Device staging of a host call: each array is copied in (outputs too, so
// that the cells the reference leaves untouched keep the caller's values)
// and the outputs copied back.
struct Staging {
    std::vector<void *> bufs;
    ~Staging() {
        for (void *p : bufs) (void)hipFree(p);
    }
    template <typename T>
    int in(const T *src, size_t n, bool dev, hipStream_t s, T **out) {
        if (!src) { *out = nullptr; return MGPU_SUCCESS; }
        if (dev) { *out = const_cast<T *>(src); return MGPU_SUCCESS; }
        void *p = nullptr;
        if (hipMalloc(&p, sizeof(T) * n) != hipSuccess) return MGPU_ERR_NOMEM;
        bufs.push_back(p);
        HIPCHK(hipMemcpyAsync(p, src, sizeof(T) * n, hipMemcpyHostToDevice, s));
        *out = static_cast<T *>(p);
        return MGPU_SUCCESS;
    }
};

int finish(hipStream_t s) {
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    return MGPU_SUCCESS;
}

}  // namespace

extern "C" int mgpu_waterflux_z(const mgpu_size3d *size, const mgpu_size3d *ws,
                                const double *Volume_Z_New, const double *Volume_Z_Old,
                                const double *WaterFlux_X, const double *WaterFlux_Y,
                                double *WaterFlux_Z, const int32_t *ComputeFaces3D_U,
                                const int32_t *ComputeFaces3D_V, const int32_t *WaterPoints3D,
                                double DT, int32_t variable_grid, int32_t device_pointers) {
    int rc = check_dims(size, ws);
    if (rc) return rc;
    if (!WaterFlux_X || !WaterFlux_Y || !WaterFlux_Z || !ComputeFaces3D_U ||
        !ComputeFaces3D_V || !WaterPoints3D || (variable_grid && (!Volume_Z_New || !Volume_Z_Old)))
        return MGPU_ERR_BADARG;
    const Dims d = make_dims(*size, *ws);
    const size_t n = (size_t)d.n3();
    const bool dev = device_pointers != 0;
    hipStream_t s = nullptr;
    Staging st;
    const double *vn = nullptr, *vo = nullptr, *qx = nullptr, *qy = nullptr;
    double *qz = nullptr;
    const int32_t *cu = nullptr, *cv = nullptr, *wp = nullptr;
    double *t = nullptr;
    int32_t *ti = nullptr;
    if ((rc = st.in(variable_grid ? Volume_Z_New : nullptr, n, dev, s, &t))) return rc;
    vn = t;
    if ((rc = st.in(variable_grid ? Volume_Z_Old : nullptr, n, dev, s, &t))) return rc;
    vo = t;
    if ((rc = st.in(WaterFlux_X, n, dev, s, &t))) return rc;
    qx = t;
    if ((rc = st.in(WaterFlux_Y, n, dev, s, &t))) return rc;
    qy = t;
    if ((rc = st.in(static_cast<const double *>(WaterFlux_Z), n, dev, s, &qz))) return rc;
    if ((rc = st.in(ComputeFaces3D_U, n, dev, s, &ti))) return rc;
    cu = ti;
    if ((rc = st.in(ComputeFaces3D_V, n, dev, s, &ti))) return rc;
    cv = ti;
    if ((rc = st.in(WaterPoints3D, n, dev, s, &ti))) return rc;
    wp = ti;
    const int ni = d.wIUB - d.wILB + 1, nj = d.wJUB - d.wJLB + 1;
    hipLaunchKernelGGL(waterflux_z_kernel, dim3((ni + 127) / 128, nj), dim3(128), 0, s, d, vn,
                       vo, qx, qy, qz, cu, cv, wp, DT, variable_grid != 0);
    HIPCHK(hipGetLastError());
    if (!dev) HIPCHK(hipMemcpyAsync(WaterFlux_Z, qz, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    return finish(s);
}

extern "C" int mgpu_compute_volumes(const mgpu_size3d *size, const mgpu_size3d *ws,
                                    const double *DWZ, const double *DUX, const double *DVY,
                                    double *VolumeZ, double *VolumeU, double *VolumeV,
                                    double *VolumeW, int32_t device_pointers) {
    int rc = check_dims(size, ws);
    if (rc) return rc;
    if (!DWZ || !DUX || !DVY || !VolumeZ || !VolumeU || !VolumeV) return MGPU_ERR_BADARG;
    const Dims d = make_dims(*size, *ws);
    const size_t n = (size_t)d.n3(), n2 = (size_t)d.n2();
    const bool dev = device_pointers != 0;
    hipStream_t s = nullptr;
    Staging st;
    double *dwz, *dux, *dvy, *vz, *vu, *vv, *vw;
    if ((rc = st.in(DWZ, n, dev, s, &dwz)) || (rc = st.in(DUX, n2, dev, s, &dux)) ||
        (rc = st.in(DVY, n2, dev, s, &dvy)) || (rc = st.in(static_cast<const double *>(VolumeZ), n, dev, s, &vz)) ||
        (rc = st.in(static_cast<const double *>(VolumeU), n, dev, s, &vu)) ||
        (rc = st.in(static_cast<const double *>(VolumeV), n, dev, s, &vv)) ||
        (rc = st.in(static_cast<const double *>(VolumeW), n, dev, s, &vw)))
        return rc;
    const dim3 g((d.wIUB - d.wILB + 1 + 127) / 128, d.wJUB - d.wJLB + 1, d.wKUB - d.wKLB + 1);
    hipLaunchKernelGGL(volume_z_kernel, g, dim3(128), 0, s, d, dwz, dux, dvy, vz);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(volume_faces_kernel, g, dim3(128), 0, s, d, vz, vu, vv, vw);
    HIPCHK(hipGetLastError());
    if (!dev) {
        HIPCHK(hipMemcpyAsync(VolumeZ, vz, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(VolumeU, vu, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(VolumeV, vv, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        if (VolumeW)
            HIPCHK(hipMemcpyAsync(VolumeW, vw, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    }
    return finish(s);
}

extern "C" int mgpu_compute_areas(const mgpu_size3d *size, const mgpu_size3d *ws,
                                  const double *DUZ, const double *DVZ, const double *DXX,
                                  const double *DYY, double *AreaU, double *AreaV,
                                  const double *SZZ, const double *Coef_U,
                                  const double *CoefX_U, const double *Coef_V,
                                  const double *CoefX_V, int32_t device_pointers) {
    int rc = check_dims(size, ws);
    if (rc) return rc;
    if (!DUZ || !DVZ || !DXX || !DYY || !AreaU || !AreaV) return MGPU_ERR_BADARG;
    if (SZZ && (!Coef_U || !CoefX_U || !Coef_V || !CoefX_V)) return MGPU_ERR_BADARG;
    const Dims d = make_dims(*size, *ws);
    const size_t n = (size_t)d.n3(), n2 = (size_t)d.n2(), nk = (size_t)(d.KUB - d.KLB + 1);
    const bool dev = device_pointers != 0;
    hipStream_t s = nullptr;
    Staging st;
    double *duz, *dvz, *dxx, *dyy, *au, *av, *szz, *cu, *cxu, *cv, *cxv;
    if ((rc = st.in(DUZ, n, dev, s, &duz)) || (rc = st.in(DVZ, n, dev, s, &dvz)) ||
        (rc = st.in(DXX, n2, dev, s, &dxx)) || (rc = st.in(DYY, n2, dev, s, &dyy)) ||
        (rc = st.in(static_cast<const double *>(AreaU), n, dev, s, &au)) ||
        (rc = st.in(static_cast<const double *>(AreaV), n, dev, s, &av)) ||
        (rc = st.in(SZZ, n, dev, s, &szz)) || (rc = st.in(Coef_U, nk, dev, s, &cu)) ||
        (rc = st.in(CoefX_U, nk, dev, s, &cxu)) || (rc = st.in(Coef_V, nk, dev, s, &cv)) ||
        (rc = st.in(CoefX_V, nk, dev, s, &cxv)))
        return rc;
    const dim3 g((d.wIUB - d.wILB + 1 + 127) / 128, d.wJUB - d.wJLB + 1, d.wKUB - d.wKLB + 1);
    hipLaunchKernelGGL(areas_kernel, g, dim3(128), 0, s, d, duz, dvz, dxx, dyy, au, av, szz, cu,
                       cxu, cv, cxv);
    HIPCHK(hipGetLastError());
    if (!dev) {
        HIPCHK(hipMemcpyAsync(AreaU, au, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(AreaV, av, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    }
    return finish(s);
}
