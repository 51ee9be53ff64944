// RSS files (rockseis::File, file.cpp:184-260) read into and written from
// device memory, and the in-memory replacement of the inversion driver's
// file stacking (Image2D::stackImage_parallel, image.cpp:533-600).
//
// The reference reads a DATA2D / DATA3D file trace by trace through an
// fstream (4 or 6 header floats, then nt samples, data.cpp:463-505).  Here
// the whole body is read with one read() into pinned host memory and one
// strided copy (hipMemcpy2DAsync: source pitch = header + samples, width =
// samples) lands the samples on the device in [ntrace][nt] order while the
// headers stay on the host; writing is the reverse.  The layout is the
// reference's, bit for bit: the magic "R0CKS", int data_format,
// header_format and type, size_t Nheader and Ndims, 9 size_t n, 9 double d,
// 9 double o, then the data.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "rsg_common.h"

namespace rsg {

namespace {

constexpr char kMagic[5] = {'R', '0', 'C', 'K', 'S'};
constexpr size_t kHeaderBytes = 5 + 3 * sizeof(int) + 2 * sizeof(size_t) + 9 * sizeof(size_t) + 18 * sizeof(double);

struct FileCloser {
    FILE *f = nullptr;
    ~FileCloser() {
        if (f) fclose(f);
    }
};

struct Pinned {
    void *p = nullptr;
    ~Pinned() {
        if (p) (void)hipHostFree(p);
    }
};

int read_header(FILE *f, rsg_rss_header *h, const char *path) {
    char magic[5];
    if (fread(magic, 1, 5, f) != 5 || memcmp(magic, kMagic, 5) != 0)
        return fail(RSG_EINVAL, std::string("rss: not an RSS file: ") + path);
    int ints[3];
    size_t sz[2];
    if (fread(ints, sizeof(int), 3, f) != 3 || fread(sz, sizeof(size_t), 2, f) != 2 ||
        fread(h->n, sizeof(size_t), 9, f) != 9 || fread(h->d, sizeof(double), 9, f) != 9 ||
        fread(h->o, sizeof(double), 9, f) != 9)
        return fail(RSG_EINVAL, std::string("rss: truncated header: ") + path);
    h->data_format = ints[0];
    h->header_format = ints[1];
    h->type = ints[2];
    h->nheader = sz[0];
    h->ndims = sz[1];
    h->start = kHeaderBytes;
    return RSG_OK;
}

int write_header(FILE *f, const rsg_rss_header *h) {
    const int ints[3] = {h->data_format, h->header_format, h->type};
    const size_t sz[2] = {h->nheader, h->ndims};
    if (fwrite(kMagic, 1, 5, f) != 5 || fwrite(ints, sizeof(int), 3, f) != 3 || fwrite(sz, sizeof(size_t), 2, f) != 2 ||
        fwrite(h->n, sizeof(size_t), 9, f) != 9 || fwrite(h->d, sizeof(double), 9, f) != 9 ||
        fwrite(h->o, sizeof(double), 9, f) != 9)
        return fail(RSG_EINVAL, "rss: header write failed");
    return RSG_OK;
}

// Traces of a DATA2D (nhead 4: sx, sz, gx, gz) or DATA3D (nhead 6) file.
int read_traces(const char *path, int type, int nhead, float *coords, float *data, int ntrace, int nt,
                hipStream_t st) {
    if (!path || !data || ntrace < 0 || nt < 1) return fail(RSG_EINVAL, "rss: bad arguments");
    FileCloser fc;
    fc.f = fopen(path, "rb");
    if (!fc.f) return fail(RSG_EINVAL, std::string("rss: cannot open ") + path);
    rsg_rss_header h;
    if (int rc = read_header(fc.f, &h, path)) return rc;
    if (h.type != type || h.nheader != (size_t)nhead || h.data_format != 4 || h.header_format != 4)
        return fail(RSG_EINVAL, std::string("rss: not a float trace file of this kind: ") + path);
    if (h.n[0] != (size_t)nt || h.n[1] != (size_t)ntrace)
        return fail(RSG_EINVAL, std::string("rss: trace count or length differs from the file's: ") + path);
    const size_t per = (size_t)nhead + nt, n = per * ntrace;
    if (n == 0) return RSG_OK;
    Pinned buf;
    RSG_HIP(hipHostMalloc(&buf.p, n * sizeof(float), hipHostMallocDefault));
    if (fread(buf.p, sizeof(float), n, fc.f) != n) return fail(RSG_EINVAL, std::string("rss: truncated data: ") + path);
    const float *b = static_cast<const float *>(buf.p);
    RSG_HIP(hipMemcpy2DAsync(data, (size_t)nt * sizeof(float), b + nhead, per * sizeof(float), (size_t)nt * sizeof(float),
                             (size_t)ntrace, hipMemcpyHostToDevice, st));
    if (coords)
        for (int i = 0; i < ntrace; i++) memcpy(coords + (size_t)i * nhead, b + (size_t)i * per, nhead * sizeof(float));
    RSG_HIP(hipStreamSynchronize(st));  // the pinned buffer is freed on return
    return RSG_OK;
}

int write_traces(const char *path, int type, int nhead, const float *coords, const float *data, int ntrace, int nt,
                 float dt, float ot, hipStream_t st) {
    if (!path || !coords || !data || ntrace < 0 || nt < 1) return fail(RSG_EINVAL, "rss: bad arguments");
    const size_t per = (size_t)nhead + nt, n = per * ntrace;
    Pinned buf;
    if (n) {
        RSG_HIP(hipHostMalloc(&buf.p, n * sizeof(float), hipHostMallocDefault));
        float *b = static_cast<float *>(buf.p);
        RSG_HIP(hipMemcpy2DAsync(b + nhead, per * sizeof(float), data, (size_t)nt * sizeof(float),
                                 (size_t)nt * sizeof(float), (size_t)ntrace, hipMemcpyDeviceToHost, st));
        for (int i = 0; i < ntrace; i++) memcpy(b + (size_t)i * per, coords + (size_t)i * nhead, nhead * sizeof(float));
        RSG_HIP(hipStreamSynchronize(st));
    }
    FileCloser fc;
    fc.f = fopen(path, "wb");
    if (!fc.f) return fail(RSG_EINVAL, std::string("rss: cannot create ") + path);
    // Data2D/3D::write (data.cpp:603-612, 1287-1296): n1 = nt, d1 = dt,
    // o1 = ot, n2 = ntrace, d2 = 1; the other dimensions stay 0 and Ndims 1
    // (File's defaults)
    rsg_rss_header h{};
    h.data_format = h.header_format = 4;
    h.type = type;
    h.nheader = (size_t)nhead;
    h.ndims = 1;
    h.n[0] = (size_t)nt;
    h.d[0] = dt;
    h.o[0] = ot;
    h.n[1] = (size_t)ntrace;
    h.d[1] = 1.0;
    if (int rc = write_header(fc.f, &h)) return rc;
    if (n && fwrite(buf.p, sizeof(float), n, fc.f) != n) return fail(RSG_EINVAL, "rss: data write failed");
    return RSG_OK;
}

// Image2D::stackImage_parallel's loop (image.cpp:586-600) on the device: one
// thread per local cell (ix, iz, ihx, ihz) inside the pads; global cells
// outside the image are skipped.
__global__ void stack_image_2d_kernel(float *__restrict__ gimg, long nxg, long nzg, const float *__restrict__ limg,
                                      long nxl, long nzl, int nhx, int nhz, int ix0, int iz0, int padlx, int padhx,
                                      int padlz, int padhz) {
    const long ix = blockIdx.x * (long)blockDim.x + threadIdx.x + padlx;
    const long iz = blockIdx.y + padlz;
    const int h = blockIdx.z, ihx = h % nhx, ihz = h / nhx;
    if (ix >= nxl - padhx || iz >= nzl - padhz) return;
    const long gx = ix + ix0, gz = iz + iz0;
    if (gx < 0 || gx >= nxg || gz < 0 || gz >= nzg) return;
    const long hoff = (long)ihz * nhx + ihx;
    gimg[(hoff * nzg + gz) * nxg + gx] += limg[(hoff * nzl + iz) * nxl + ix];
}

}  // namespace

}  // namespace rsg

extern "C" {

int rsg_rss_read_header(const char *path, rsg_rss_header *h) {
    if (!path || !h) return rsg::fail(RSG_EINVAL, "rsg_rss_read_header: null argument");
    rsg::FileCloser fc;
    fc.f = fopen(path, "rb");
    if (!fc.f) return rsg::fail(RSG_EINVAL, std::string("rsg_rss_read_header: cannot open ") + path);
    return rsg::read_header(fc.f, h, path);
}

int rsg_rss_read_data2d(const char *path, float *coords, float *data, int ntrace, int nt, void *stream) {
    return rsg::read_traces(path, RSG_RSS_DATA2D, 4, coords, data, ntrace, nt, (hipStream_t)stream);
}

int rsg_rss_read_data3d(const char *path, float *coords, float *data, int ntrace, int nt, void *stream) {
    return rsg::read_traces(path, RSG_RSS_DATA3D, 6, coords, data, ntrace, nt, (hipStream_t)stream);
}

int rsg_rss_write_data2d(const char *path, const float *coords, const float *data, int ntrace, int nt, float dt,
                         float ot, void *stream) {
    return rsg::write_traces(path, RSG_RSS_DATA2D, 4, coords, data, ntrace, nt, dt, ot, (hipStream_t)stream);
}

// Data3D::write records the time origin as 0 (data.cpp:1289).
int rsg_rss_write_data3d(const char *path, const float *coords, const float *data, int ntrace, int nt, float dt,
                         void *stream) {
    return rsg::write_traces(path, RSG_RSS_DATA3D, 6, coords, data, ntrace, nt, dt, 0.0f, (hipStream_t)stream);
}

int rsg_rss_read_regular(const char *path, float *data, size_t n, void *stream) {
    if (!path || !data) return rsg::fail(RSG_EINVAL, "rsg_rss_read_regular: null argument");
    rsg::FileCloser fc;
    fc.f = fopen(path, "rb");
    if (!fc.f) return rsg::fail(RSG_EINVAL, std::string("rsg_rss_read_regular: cannot open ") + path);
    rsg_rss_header h;
    if (int rc = rsg::read_header(fc.f, &h, path)) return rc;
    size_t ntot = 1;  // Geometry::getNtot: the product of the non-zero sizes
    for (int i = 0; i < 9; i++) ntot *= h.n[i] ? h.n[i] : 1;
    if (h.nheader != 0 || h.data_format != 4 || ntot != n)
        return rsg::fail(RSG_EINVAL, std::string("rsg_rss_read_regular: not a float array of that size: ") + path);
    if (n == 0) return RSG_OK;
    rsg::Pinned buf;
    RSG_HIP(hipHostMalloc(&buf.p, n * sizeof(float), hipHostMallocDefault));
    if (fread(buf.p, sizeof(float), n, fc.f) != n)
        return rsg::fail(RSG_EINVAL, std::string("rsg_rss_read_regular: truncated data: ") + path);
    RSG_HIP(hipMemcpyAsync(data, buf.p, n * sizeof(float), hipMemcpyHostToDevice, (hipStream_t)stream));
    RSG_HIP(hipStreamSynchronize((hipStream_t)stream));
    return RSG_OK;
}

int rsg_rss_write_regular(const char *path, const rsg_rss_header *h, const float *data, void *stream) {
    if (!path || !h || !data) return rsg::fail(RSG_EINVAL, "rsg_rss_write_regular: null argument");
    if (h->nheader != 0 || h->data_format != 4)
        return rsg::fail(RSG_EINVAL, "rsg_rss_write_regular: float arrays without trace headers only");
    size_t n = 1;
    for (int i = 0; i < 9; i++) n *= h->n[i] ? h->n[i] : 1;
    rsg::Pinned buf;
    if (n) {
        RSG_HIP(hipHostMalloc(&buf.p, n * sizeof(float), hipHostMallocDefault));
        RSG_HIP(hipMemcpyAsync(buf.p, data, n * sizeof(float), hipMemcpyDeviceToHost, (hipStream_t)stream));
        RSG_HIP(hipStreamSynchronize((hipStream_t)stream));
    }
    rsg::FileCloser fc;
    fc.f = fopen(path, "wb");
    if (!fc.f) return rsg::fail(RSG_EINVAL, std::string("rsg_rss_write_regular: cannot create ") + path);
    if (int rc = rsg::write_header(fc.f, h)) return rc;
    if (n && fwrite(buf.p, sizeof(float), n, fc.f) != n)
        return rsg::fail(RSG_EINVAL, "rsg_rss_write_regular: data write failed");
    return RSG_OK;
}

int rsg_stack_image_2d(float *gimg, const rsg_image2d *g, const float *limg, const rsg_image2d *l, int padlx,
                       int padhx, int padlz, int padhz, void *stream) {
    if (!gimg || !g || !limg || !l) return rsg::fail(RSG_EINVAL, "rsg_stack_image_2d: null argument");
    if (g->nhx != l->nhx || g->nhz != l->nhz || g->dx != l->dx || g->dz != l->dz)
        return rsg::fail(RSG_EINVAL, "rsg_stack_image_2d: images are not compatible, cannot stack");
    if (l->nx < 1 || l->nz < 1 || l->nhx < 1 || l->nhz < 1 || g->nx < 1 || g->nz < 1 || l->nz > 65535 ||
        (long)l->nhx * l->nhz > 65535 || padlx < 0 || padhx < 0 || padlz < 0 || padhz < 0)
        return rsg::fail(RSG_EINVAL, "rsg_stack_image_2d: bad sizes or pads");
    const int ix0 = rsg_image_offset(l->ox, l->dx, g->ox, g->dx), iz0 = rsg_image_offset(l->oz, l->dz, g->oz, g->dz);
    const long nxs = (long)l->nx - padlx - padhx, nzs = (long)l->nz - padlz - padhz;
    if (nxs <= 0 || nzs <= 0) return RSG_OK;
    const dim3 grid((unsigned)((nxs + 255) / 256), (unsigned)nzs, (unsigned)(l->nhx * l->nhz));
    rsg::stack_image_2d_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(gimg, g->nx, g->nz, limg, l->nx, l->nz, l->nhx,
                                                                       l->nhz, ix0, iz0, padlx, padhx, padlz, padhz);
    RSG_LAUNCH_CHECK("stack_image_2d_kernel");
    return RSG_OK;
}

}  // extern "C"
