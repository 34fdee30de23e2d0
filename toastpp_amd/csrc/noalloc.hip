// ProjectAll on GPU-resident fields without any device allocation.
//
// tmi_solver_project_all_device stages its inputs in buffers it allocates
// per call: the (detector, source) -> link table, the measurement vectors'
// CSR arrays and the projections before their copy to the host.  Callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc) use the variant below,
// which stages the same data in one file-scope device array, `scratch`, of
// TMI_NOALLOC_SCRATCH_BYTES, and runs the same projection kernel.
//
// The array exists once per GPU for the life of the process; its device
// address is looked up once per device (hipGetSymbolAddress, with that device
// current) and only that address is handed to copies and kernels -- on the
// host the array's name is a stand-in, not a device address.  Each call
// overwrites the whole region it uses (the link table is written entry by
// entry, every linked projection by the kernel), so nothing needs clearing;
// calls on one GPU share the array, and a lock serialises them within the
// process (calls from two processes on one GPU each have their own copy).
#include <hip/hip_runtime.h>

#include <complex>
#include <mutex>

#include "handles.hpp"

namespace tmi {

void set_last_error(const std::string &msg);

static __device__ __align__(256) unsigned char scratch[TMI_NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static void *g_scratch[NOALLOC_MAX_DEVICES];
static std::once_flag g_scratch_once[NOALLOC_MAX_DEVICES];
static int g_scratch_status[NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;

// device address of this GPU's copy of `scratch` (the current device's)
static unsigned char *scratch_address()
{
    int dev = 0;
    TMI_HIP(hipGetDevice(&dev));
    TMI_CHECK(dev >= 0 && dev < NOALLOC_MAX_DEVICES, TMI_ERR_DEVICE, "device %d beyond the supported range", dev);
    std::call_once(g_scratch_once[dev], [dev] {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch_status[dev] = (int)e;
        if (e == hipSuccess) g_scratch[dev] = p;
    });
    TMI_CHECK(g_scratch_status[dev] == (int)hipSuccess && g_scratch[dev], TMI_ERR_DEVICE,
              "hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
              hipGetErrorString((hipError_t)g_scratch_status[dev]));
    return (unsigned char *)g_scratch[dev];
}

// link table of the projection kernel: lk[m * nq + q] = index of the link
// (q, m) in the source-major link list, -1 where q does not see m.  One
// thread per (detector, source) pair scans the source's links.
__global__ void k_link_table(int nq, int nm, const int *__restrict__ lptr, const int *__restrict__ lidx,
                             int *__restrict__ lk)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nq * nm) return;
    const int m = (int)(t / nq), q = (int)(t % nq);
    int k = -1;
    for (int j = lptr[q]; j < lptr[q + 1]; j++)
        if (lidx[j] == m) { k = j; break; }
    lk[t] = k;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace tmi

using namespace tmi;

extern "C" int tmi_solver_project_all_device_noalloc(tmi_solver h, int nm, const int *mrp, const int *mci,
                                                     const double *mval, const void *d_phi, int ld, int dscale,
                                                     double *proj)
{
    try {
        TMI_CHECK(h && mrp && mci && mval && d_phi && proj, TMI_ERR_INVALID_ARG, "NULL argument");
        Solver &s = h->s;
        Mesh &m = *s.mesh;
        const QM &qm = m.qm;
        TMI_CHECK(nm == qm.nm, TMI_ERR_INVALID_ARG, "mvec has %d rows, the QM description %d detectors", nm, qm.nm);
        TMI_CHECK(ld >= qm.nq, TMI_ERR_INVALID_ARG, "field row stride %d below the source count %d", ld, qm.nq);
        TMI_CHECK(dscale == TMI_DATA_LIN || dscale == TMI_DATA_LOG, TMI_ERR_INVALID_ARG, "unknown data scale %d",
                  dscale);
        const int nq = qm.nq, nqm = qm.nqm, cw = s.cplx ? 2 : 1;
        if (nqm == 0) return TMI_OK;
        TMI_CHECK(mrp[0] == 0 && mrp[nm] >= 0, TMI_ERR_INVALID_ARG, "invalid mvec row pointer");
        int maxrow = 0;
        for (int i = 0; i < nm; i++) {
            TMI_CHECK(mrp[i + 1] >= mrp[i], TMI_ERR_INVALID_ARG, "mvec row pointer not ascending");
            maxrow = std::max(maxrow, mrp[i + 1] - mrp[i]);
        }
        const long mnnz = mrp[nm];
        const ProjectPlan pl = project_plan(nq, nm, maxrow, s.cplx ? 1 : 0);
        // scratch layout (each part 256-byte aligned): link list pointer and
        // indices, link table, mvec row pointer, columns, values, segment
        // partial sums, projections.
        // The size check comes before any read of the CSR arrays.
        const size_t o_lp = 0, o_li = o_lp + align256(sizeof(int) * (nq + 1)),
                     o_lk = o_li + align256(sizeof(int) * (size_t)nqm),
                     o_rp = o_lk + align256(sizeof(int) * (size_t)nq * nm),
                     o_ci = o_rp + align256(sizeof(int) * (size_t)(nm + 1)),
                     o_val = o_ci + align256(sizeof(int) * (size_t)mnnz),
                     o_part = o_val + align256(sizeof(double) * (size_t)mnnz * cw),
                     o_pr = o_part + align256(sizeof(double) * pl.partial_doubles),
                     need = o_pr + align256(sizeof(double) * (size_t)nqm * cw);
        TMI_CHECK(need <= (size_t)TMI_NOALLOC_SCRATCH_BYTES, TMI_ERR_INVALID_ARG,
                  "ProjectAll without allocation needs %zu bytes of scratch for %d sources, %d detectors and %ld mvec "
                  "entries; the library's scratch holds %zu",
                  need, nq, nm, mnnz, (size_t)TMI_NOALLOC_SCRATCH_BYTES);
        for (long k = 0; k < mnnz; k++)
            TMI_CHECK(mci[k] >= 0 && mci[k] < m.n, TMI_ERR_INVALID_ARG, "mvec column out of range");
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        unsigned char *base = scratch_address();
        hipStream_t st = s.st;
        int *d_lp = (int *)(base + o_lp), *d_li = (int *)(base + o_li), *d_lk = (int *)(base + o_lk);
        int *d_rp = (int *)(base + o_rp), *d_ci = (int *)(base + o_ci);
        double *d_val = (double *)(base + o_val), *d_part = (double *)(base + o_part), *d_pr = (double *)(base + o_pr);
        TMI_HIP(hipMemcpyAsync(d_lp, qm.lptr.data(), sizeof(int) * (nq + 1), hipMemcpyHostToDevice, st));
        TMI_HIP(hipMemcpyAsync(d_li, qm.lidx.data(), sizeof(int) * (size_t)nqm, hipMemcpyHostToDevice, st));
        TMI_HIP(hipMemcpyAsync(d_rp, mrp, sizeof(int) * (size_t)(nm + 1), hipMemcpyHostToDevice, st));
        if (mnnz) {
            TMI_HIP(hipMemcpyAsync(d_ci, mci, sizeof(int) * (size_t)mnnz, hipMemcpyHostToDevice, st));
            TMI_HIP(hipMemcpyAsync(d_val, mval, sizeof(double) * (size_t)mnnz * cw, hipMemcpyHostToDevice, st));
        }
        const long pairs = (long)nq * nm;
        hipLaunchKernelGGL(k_link_table, dim3(grid_for(pairs, 256)), dim3(256), 0, st, nq, nm, d_lp, d_li, d_lk);
        TMI_LAUNCH_CHECK();
        launch_project(pl, nq, nm, d_lk, d_rp, d_ci, d_val, (const double *)d_phi, 1, ld, s.cplx ? 1 : 0, d_part,
                       d_pr, st);
        TMI_HIP(hipMemcpyAsync(proj, d_pr, sizeof(double) * (size_t)nqm * cw, hipMemcpyDeviceToHost, st));
        TMI_HIP(hipStreamSynchronize(st));
        if (dscale == TMI_DATA_LOG) {   // complex log: Re = ln|y|, Im = arg y
            for (int k = 0; k < nqm; k++) {
                if (s.cplx) {
                    const std::complex<double> z = std::log(std::complex<double>(proj[2 * k], proj[2 * k + 1]));
                    proj[2 * k] = z.real();
                    proj[2 * k + 1] = z.imag();
                } else {
                    proj[k] = std::log(proj[k]);
                }
            }
        }
        return TMI_OK;
    } catch (const Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return TMI_ERR_INVALID_ARG;
    }
}
