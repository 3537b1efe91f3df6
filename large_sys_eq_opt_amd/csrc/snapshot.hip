// snapshot.hip -- snapshot output at speed: a device copy of the state taken
// on the caller's stream, a device-to-host copy of it into pinned host
// memory on a stream of its own (stepping goes on meanwhile), and the
// reference's CSV format written by several host threads.
//
// The reference downloads the state with a blocking copy between steps and
// formats it cell by cell on one thread (scripts/gpu/main_gpu.cu:272-298,
// :146-171); its snapshot writes took 1.22 s of a 2.16 s GPU run
// (outputs/baseline/gpu_timing.txt:21-27).
//
// CSV: header "i,j,x,y,rho,rhou,rhov,E", one row per interior cell in the
// order j = ng..ng+ny-1, i = ng..ng+nx-1 (indices include the ghost
// offset), x = x_min + ((i - ng) + 0.5) dx, y likewise, values printed as
// std::ostream prints a double at the given precision (%.{p}g; the reference
// uses the default 6): scripts/gpu/main_gpu.cu:146-171,
// scripts/cpu/main_cpu.cpp:136-164.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <charconv>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "grid.h"

namespace mh {
namespace {

// Appends "%.{prec}g" of v (what an ostream with that precision prints):
// std::to_chars in general format with a precision is specified to give
// printf's %.{prec}g output, without printf's per-call format parsing.
inline void put_g(std::string& o, double v, int prec) {
    char b[40];
    const std::to_chars_result r = std::to_chars(b, b + sizeof b, v, std::chars_format::general, prec);
    o.append(b, (size_t)(r.ptr - b));
}
inline void put_i(std::string& o, int v) {
    char b[16];
    const std::to_chars_result r = std::to_chars(b, b + sizeof b, v);
    o.append(b, (size_t)(r.ptr - b));
}

// Cell (i, j) of the (ny+4) x (nx+4) ghosted array, conserved component c.
struct HostState {
    const double* aos = nullptr;     // reference AoS layout, or
    const double* soa[4] = {};       // four planes
    int64_t tnx = 0;
    double at(int i, int j, int c) const {
        const int64_t k = (int64_t)j * tnx + i;
        return aos ? aos[4 * k + c] : soa[c][k];
    }
};

// Formats interior rows [j0, j1) (ghosted indices) into o.
void format_rows(const HostState& h, int nx, double x_min, double y_min, double dx, double dy, int prec, int j0,
                 int j1, std::string& o) {
    for (int j = j0; j < j1; ++j) {
        const double y = y_min + (static_cast<double>(j - kNg) + 0.5) * dy;
        for (int i = kNg; i < kNg + nx; ++i) {
            const double x = x_min + (static_cast<double>(i - kNg) + 0.5) * dx;
            put_i(o, i); o += ',';
            put_i(o, j); o += ',';
            put_g(o, x, prec); o += ',';
            put_g(o, y, prec);
            for (int c = 0; c < 4; ++c) {
                o += ',';
                put_g(o, h.at(i, j, c), prec);
            }
            o += '\n';
        }
    }
}

// Rows are formatted in rounds of nthreads blocks (bounded memory: about
// 70 bytes per cell of one round) and written in order.
int write_csv(const HostState& h, int nx, int ny, double x_min, double y_min, double dx, double dy, int prec,
              int nthreads, const char* path) {
    if (!path || nx < 1 || ny < 1 || prec < 1 || prec > 17)
        return fail(MH_ERR_INVALID, "write csv: bad argument");
    if (nthreads <= 0) nthreads = (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(MH_ERR_INVALID, std::string("write csv: cannot open ") + path);
    static const char head[] = "i,j,x,y,rho,rhou,rhov,E\n";
    bool ok = std::fwrite(head, 1, sizeof head - 1, f) == sizeof head - 1;
    // rows per block: about 2^16 cells
    const int rows = std::max(1, (1 << 16) / nx);
    std::vector<std::string> buf((size_t)nthreads);
    for (int j = kNg; j < kNg + ny && ok;) {
        std::vector<std::thread> pool;
        int nb = 0;
        for (int t = 0; t < nthreads && j + t * rows < kNg + ny; ++t) ++nb;
        for (int t = 0; t < nb; ++t) {
            const int a = j + t * rows, b = std::min(kNg + ny, a + rows);
            buf[t].clear();
            pool.emplace_back([&, a, b, t] { format_rows(h, nx, x_min, y_min, dx, dy, prec, a, b, buf[t]); });
        }
        for (auto& th : pool) th.join();
        for (int t = 0; t < nb && ok; ++t) ok = std::fwrite(buf[t].data(), 1, buf[t].size(), f) == buf[t].size();
        j = std::min(kNg + ny, j + nb * rows);
    }
    if (std::fclose(f) != 0) ok = false;
    if (!ok) return fail(MH_ERR_INVALID, std::string("write csv: write failed: ") + path);
    return MH_OK;
}

}  // namespace
}  // namespace mh

using namespace mh;

struct mh_snapshot_s {
    mh_grid_desc d{};
    double dx = 0, dy = 0;
    int nbuf = 0;
    std::vector<mh_grid_s*> dev;     // device copies, one per slot
    std::vector<double*> host;       // pinned SoA planes, (ny+4)*(nx+4) each, one block per slot
    std::vector<hipEvent_t> copied, landed;
    std::vector<int> pending;        // a capture not yet waited for
    hipStream_t copy = nullptr;
};

extern "C" {

int mh_snapshot_destroy(mh_snapshot s) {
    if (!s) return MH_OK;
    if (s->copy) (void)hipStreamSynchronize(s->copy);
    for (mh_grid_s* g : s->dev)
        if (g) mh_grid_destroy(g);
    for (double* p : s->host)
        if (p) (void)hipHostFree(p);
    for (hipEvent_t e : s->copied)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : s->landed)
        if (e) (void)hipEventDestroy(e);
    if (s->copy) (void)hipStreamDestroy(s->copy);
    delete s;
    return MH_OK;
}

int mh_snapshot_create(mh_grid like, int nbuf, mh_snapshot* out) {
    if (!like || !out || nbuf < 1 || nbuf > 8) return fail(MH_ERR_INVALID, "mh_snapshot_create: bad argument");
    mh_snapshot s = new (std::nothrow) mh_snapshot_s();
    if (!s) return fail(MH_ERR_OOM, "host allocation failed");
    s->d = like->d;
    s->dx = like->dx;
    s->dy = like->dy;
    s->nbuf = nbuf;
    s->dev.assign(nbuf, nullptr);
    s->host.assign(nbuf, nullptr);
    s->copied.assign(nbuf, nullptr);
    s->landed.assign(nbuf, nullptr);
    s->pending.assign(nbuf, 0);
    const size_t cells = (size_t)(s->d.nx + 4) * (size_t)(s->d.ny + 4);
    int rc = check_hip(hipStreamCreateWithFlags(&s->copy, hipStreamNonBlocking), "hipStreamCreate");
    for (int k = 0; k < nbuf && !rc; ++k) {
        rc = mh_grid_create(&s->d, &s->dev[k]);
        if (!rc) rc = check_hip(hipHostMalloc((void**)&s->host[k], 4 * cells * sizeof(double)), "hipHostMalloc(snapshot)");
        if (!rc) rc = check_hip(hipEventCreateWithFlags(&s->copied[k], hipEventDisableTiming), "hipEventCreate");
        if (!rc) rc = check_hip(hipEventCreateWithFlags(&s->landed[k], hipEventDisableTiming), "hipEventCreate");
    }
    if (rc) {
        mh_snapshot_destroy(s);
        return rc;
    }
    *out = s;
    return MH_OK;
}

int mh_snapshot_capture(mh_snapshot s, mh_grid U, int slot, mh_stream stream) {
    if (!s || !U || slot < 0 || slot >= s->nbuf) return fail(MH_ERR_INVALID, "mh_snapshot_capture: bad argument");
    if (U->d.nx != s->d.nx || U->d.ny != s->d.ny) return fail(MH_ERR_INVALID, "mh_snapshot_capture: shape mismatch");
    hipStream_t st = (hipStream_t)stream;
    mh_grid_s* g = s->dev[slot];
    // the slot's device copy may still be on its way to the host
    int rc = check_hip(hipStreamWaitEvent(st, s->landed[slot], 0), "hipStreamWaitEvent");
    if (!rc) rc = check_hip(hipMemcpyAsync(g->base, U->base, (size_t)(4 * U->ps) * sizeof(double),
                                           hipMemcpyDeviceToDevice, st),
                            "hipMemcpyAsync(snapshot)");
    if (!rc) rc = check_hip(hipEventRecord(s->copied[slot], st), "hipEventRecord");
    if (!rc) rc = check_hip(hipStreamWaitEvent(s->copy, s->copied[slot], 0), "hipStreamWaitEvent");
    const size_t w = (size_t)(s->d.nx + 4), h = (size_t)(s->d.ny + 4);
    for (int c = 0; c < 4 && !rc; ++c) {
        const double* src = g->base + (int64_t)c * g->ps + (kLead - kNg);
        rc = check_hip(hipMemcpy2DAsync(s->host[slot] + c * w * h, w * sizeof(double), src, g->rp * sizeof(double),
                                        w * sizeof(double), h, hipMemcpyDeviceToHost, s->copy),
                       "hipMemcpy2DAsync(snapshot)");
    }
    if (!rc) rc = check_hip(hipEventRecord(s->landed[slot], s->copy), "hipEventRecord");
    if (!rc) s->pending[slot] = 1;
    return rc;
}

int mh_snapshot_wait(mh_snapshot s, int slot, double** planes) {
    if (!s || slot < 0 || slot >= s->nbuf) return fail(MH_ERR_INVALID, "mh_snapshot_wait: bad argument");
    int rc = check_hip(hipEventSynchronize(s->landed[slot]), "hipEventSynchronize");
    if (rc) return rc;
    s->pending[slot] = 0;
    if (planes) {
        const size_t n = (size_t)(s->d.nx + 4) * (size_t)(s->d.ny + 4);
        for (int c = 0; c < 4; ++c) planes[c] = s->host[slot] + c * n;
    }
    return MH_OK;
}

int mh_snapshot_write_csv(mh_snapshot s, int slot, const char* path, int precision, int nthreads) {
    if (!s || slot < 0 || slot >= s->nbuf) return fail(MH_ERR_INVALID, "mh_snapshot_write_csv: bad argument");
    double* planes[4];
    int rc = mh_snapshot_wait(s, slot, planes);
    if (rc) return rc;
    HostState h;
    for (int c = 0; c < 4; ++c) h.soa[c] = planes[c];
    h.tnx = s->d.nx + 4;
    return write_csv(h, s->d.nx, s->d.ny, s->d.x_min, s->d.y_min, s->dx, s->dy, precision, nthreads, path);
}

int mh_write_csv_aos(const double* host_aos, const mh_grid_desc* desc, int precision, int nthreads, const char* path) {
    if (!host_aos || !desc || desc->nx < 1 || desc->ny < 1) return fail(MH_ERR_INVALID, "mh_write_csv_aos: bad argument");
    HostState h;
    h.aos = host_aos;
    h.tnx = desc->nx + 4;
    const double dx = (desc->x_max - desc->x_min) / (double)desc->nx;   // head/gpu/grid_gpu.cuh:40-41
    const double dy = (desc->y_max - desc->y_min) / (double)desc->ny;
    return write_csv(h, desc->nx, desc->ny, desc->x_min, desc->y_min, dx, dy, precision, nthreads, path);
}

}  // extern "C"
