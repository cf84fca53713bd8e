// Files for ParaView and numpy, written from a cache's device columns through the C ABI alone.
//
//   cc_write_cache_vtk: `prefix_cells.vti` -- VTK XML ImageData over the grid, with every cell's
//     volume fraction, kind, ambiguous flag, centroid, interface centroid and one array per face
//     direction of face fractions (the layout write_cache_vtk uses, src/tri_clipping/surface.jl:214-252),
//     phi as point data for the nodal methods and the planes for PLIC -- and, with `surface`,
//     `prefix_surface.vtu`, the cache's generate_mesh result (lines in 2D, triangles in 3D).
//   cc_write_npy: one column as a .npy array in the reference's own layout (column-major, components
//     first: an SVector field of n cells reads as shape (components, n0, n1[, n2]), fortran_order).
//
// Data are raw little-endian appended blocks, each led by its UInt64 byte count.  Columns stream
// through a host buffer of at most CHUNK elements; a component of an interleaved vector column is
// gathered on the device first, so no column is ever held whole on the host.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cutcell_hip.h"
#include "cc_internal.h"


namespace {

constexpr int64_t CHUNK = 1 << 24;  // elements per host round trip

__global__ void k_gather(const char *__restrict__ src, int64_t stride, int64_t es, int64_t n,
                         char *__restrict__ dst) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    for (int64_t b = 0; b < es; ++b) dst[q * es + b] = src[q * stride + b];
}

int io_fail(int code, const std::string &msg) {
    cc_internal_set_error(msg.c_str());
    return code;
}

struct Body {
    int method = 0, dtype = 0, kz = 0, dim = 0;
    bool leaves = false;
    int64_t n[3] = {1, 1, 1};
    void *cols[CC_NUM_FIELDS] = {};
    int64_t ncell = 1, nnode = 1;
    size_t es = 8;
};

int body_of(cc_cache *c, Body &b) {
    void *cl = nullptr, *ct = nullptr;
    if (cc_internal_body(c, &b.method, &b.dtype, b.n, b.cols, &cl, &ct, &b.kz) != CC_OK)
        return io_fail(CC_ERR_ARGUMENT, "cache is NULL");
    b.dim = b.n[2] == 0 ? 2 : 3;
    b.leaves = cc_internal_leaves(c) != 0;
    b.ncell = b.nnode = 1;
    for (int a = 0; a < b.dim; ++a) {
        b.ncell *= b.n[a];
        b.nnode *= b.n[a] + 1;
    }
    b.es = b.dtype == CC_FLOAT32 ? 4 : 8;
    return CC_OK;
}

// Streams component `comp` of an interleaved column (`ncomp` values of `es` bytes per entry, `count`
// entries) to f, `chunk` entries per round trip (host and stage hold chunk * es bytes), converted by
// `convert` (bytes in, bytes out) when given.
int stream_column(FILE *f, const void *dev, int64_t count, int ncomp, int comp, size_t es, hipStream_t st,
                  char *stage, std::vector<char> &host, int64_t chunk, size_t out_es = 0,
                  void (*convert)(const char *, char *, int64_t) = nullptr) {
    for (int64_t q0 = 0; q0 < count; q0 += chunk) {
        const int64_t m = count - q0 < chunk ? count - q0 : chunk;
        const char *src = (const char *)dev + (size_t)(q0 * ncomp + comp) * es;
        if (ncomp == 1) {
            if (hipMemcpyAsync(host.data(), src, m * es, hipMemcpyDeviceToHost, st) != hipSuccess)
                return io_fail(CC_ERR_HIP, "copy to host failed");
        } else {
            hipLaunchKernelGGL(k_gather, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, src,
                               (int64_t)(ncomp * es), (int64_t)es, m, stage);
            if (hipMemcpyAsync(host.data(), stage, m * es, hipMemcpyDeviceToHost, st) != hipSuccess)
                return io_fail(CC_ERR_HIP, "copy to host failed");
        }
        if (hipStreamSynchronize(st) != hipSuccess) return io_fail(CC_ERR_HIP, "copy to host failed");
        if (convert) {
            std::vector<char> out(m * out_es);
            convert(host.data(), out.data(), m);
            if (fwrite(out.data(), 1, out.size(), f) != out.size()) return io_fail(CC_ERR_ARGUMENT, "write failed");
        } else if (fwrite(host.data(), 1, m * es, f) != (size_t)(m * es)) {
            return io_fail(CC_ERR_ARGUMENT, "write failed");
        }
    }
    return CC_OK;
}

void i8_to_i32(const char *in, char *out, int64_t m) {
    for (int64_t q = 0; q < m; ++q) {
        const int32_t v = (int8_t)in[q];
        std::memcpy(out + 4 * q, &v, 4);
    }
}
void u8_to_i32(const char *in, char *out, int64_t m) {
    for (int64_t q = 0; q < m; ++q) {
        const int32_t v = (uint8_t)in[q];
        std::memcpy(out + 4 * q, &v, 4);
    }
}

struct Array {
    std::string name, type;
    int ncomp;       // components written
    int field;       // cache column
    int stride;      // values per entry in the column
    int comp;        // first component within an entry (a single component when ncomp == 1 < stride)
    size_t in_es, out_es;
    void (*convert)(const char *, char *, int64_t);
    int64_t count;   // entries
    bool point;      // point data (phi) rather than cell data
};

int write_vti(const char *path, const Body &b, const cc_grid *g, const std::vector<Array> &arrays, hipStream_t st) {
    FILE *f = fopen(path, "wb");
    if (!f) return io_fail(CC_ERR_ARGUMENT, std::string("cannot open ") + path);
    std::string xml = "<?xml version=\"1.0\"?>\n<VTKFile type=\"ImageData\" version=\"1.0\" "
                      "byte_order=\"LittleEndian\" header_type=\"UInt64\">\n";
    char buf[512];
    snprintf(buf, sizeof buf, "  <ImageData WholeExtent=\"0 %lld 0 %lld 0 %lld\" Origin=\"%.17g %.17g %.17g\" "
             "Spacing=\"%.17g %.17g %.17g\">\n    <Piece Extent=\"0 %lld 0 %lld 0 %lld\">\n",
             (long long)b.n[0], (long long)b.n[1], (long long)(b.dim == 3 ? b.n[2] : 0), g->x0[0], g->x0[1],
             b.dim == 3 ? g->x0[2] : 0.0, g->d[0], g->d[1], b.dim == 3 ? g->d[2] : 1.0, (long long)b.n[0],
             (long long)b.n[1], (long long)(b.dim == 3 ? b.n[2] : 0));
    xml += buf;
    uint64_t offset = 0;
    for (int pass = 0; pass < 2; ++pass) {
        bool any = false;
        for (const Array &a : arrays)
            if (a.point == (pass == 0)) any = true;
        if (!any) continue;
        xml += pass == 0 ? "      <PointData>\n" : "      <CellData>\n";
        for (const Array &a : arrays) {
            if (a.point != (pass == 0)) continue;
            snprintf(buf, sizeof buf,
                     "        <DataArray type=\"%s\" Name=\"%s\" NumberOfComponents=\"%d\" format=\"appended\" "
                     "offset=\"%llu\"/>\n",
                     a.type.c_str(), a.name.c_str(), a.ncomp, (unsigned long long)offset);
            xml += buf;
            offset += 8 + (uint64_t)a.count * a.ncomp * a.out_es;
        }
        xml += pass == 0 ? "      </PointData>\n" : "      </CellData>\n";
    }
    xml += "    </Piece>\n  </ImageData>\n  <AppendedData encoding=\"raw\">\n   _";
    if (fwrite(xml.data(), 1, xml.size(), f) != xml.size()) { fclose(f); return io_fail(CC_ERR_ARGUMENT, "write failed"); }
    // the largest single round trip: every column's entries, at most CHUNK of them
    int64_t chunk = 1;
    for (const Array &a : arrays) {
        const int64_t e = a.ncomp == a.stride ? a.count * a.stride : a.count;
        chunk = e > chunk ? e : chunk;
    }
    chunk = chunk < CHUNK ? chunk : CHUNK;
    char *stage = nullptr;
    int rc = CC_OK;
    std::vector<char> host((size_t)chunk * 8);
    if (hipMalloc(&stage, (size_t)chunk * 8) != hipSuccess) { fclose(f); return io_fail(CC_ERR_OUT_OF_MEMORY, "staging"); }
    for (int pass = 0; pass < 2 && rc == CC_OK; ++pass)
        for (const Array &a : arrays) {
            if (a.point != (pass == 0)) continue;
            const uint64_t nbytes = (uint64_t)a.count * a.ncomp * a.out_es;
            if (fwrite(&nbytes, 8, 1, f) != 1) {
                rc = io_fail(CC_ERR_ARGUMENT, "write failed");
                break;
            }
            if (a.ncomp == a.stride) {  // a whole column, entry by entry
                rc = stream_column(f, b.cols[a.field], a.count * a.stride, 1, 0, a.in_es, st, stage, host, chunk,
                                   a.out_es, a.convert);
            } else {  // one component of an interleaved column
                rc = stream_column(f, b.cols[a.field], a.count, a.stride, a.comp, a.in_es, st, stage, host, chunk,
                                   a.out_es, a.convert);
            }
            if (rc) break;
        }
    (void)hipFree(stage);
    const char *tail = "\n  </AppendedData>\n</VTKFile>\n";
    if (rc == CC_OK && fwrite(tail, 1, strlen(tail), f) != strlen(tail)) rc = io_fail(CC_ERR_ARGUMENT, "write failed");
    fclose(f);
    return rc;
}

int write_vtu(const char *path, int dim, int dtype, const std::vector<char> &nodes, int64_t nn,
              const std::vector<int64_t> &elems, int64_t ne) {
    FILE *f = fopen(path, "wb");
    if (!f) return io_fail(CC_ERR_ARGUMENT, std::string("cannot open ") + path);
    const size_t es = dtype == CC_FLOAT32 ? 4 : 8;
    const int per = dim;  // 2 nodes per line, 3 per triangle
    std::string xml = "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" "
                      "byte_order=\"LittleEndian\" header_type=\"UInt64\">\n  <UnstructuredGrid>\n";
    char buf[512];
    snprintf(buf, sizeof buf, "    <Piece NumberOfPoints=\"%lld\" NumberOfCells=\"%lld\">\n", (long long)nn,
             (long long)ne);
    xml += buf;
    const uint64_t pbytes = (uint64_t)nn * 3 * es, cbytes = (uint64_t)ne * per * 8, obytes = (uint64_t)ne * 8,
                   tbytes = (uint64_t)ne;
    snprintf(buf, sizeof buf,
             "      <Points>\n        <DataArray type=\"%s\" NumberOfComponents=\"3\" format=\"appended\" "
             "offset=\"0\"/>\n      </Points>\n      <Cells>\n"
             "        <DataArray type=\"Int64\" Name=\"connectivity\" format=\"appended\" offset=\"%llu\"/>\n"
             "        <DataArray type=\"Int64\" Name=\"offsets\" format=\"appended\" offset=\"%llu\"/>\n"
             "        <DataArray type=\"UInt8\" Name=\"types\" format=\"appended\" offset=\"%llu\"/>\n"
             "      </Cells>\n",
             es == 4 ? "Float32" : "Float64", (unsigned long long)(8 + pbytes),
             (unsigned long long)(16 + pbytes + cbytes), (unsigned long long)(24 + pbytes + cbytes + obytes));
    xml += buf;
    xml += "    </Piece>\n  </UnstructuredGrid>\n  <AppendedData encoding=\"raw\">\n   _";
    bool ok = fwrite(xml.data(), 1, xml.size(), f) == xml.size();
    // points, padded to three coordinates in 2D
    std::vector<char> pts(pbytes);
    for (int64_t q = 0; q < nn; ++q)
        for (int x = 0; x < 3; ++x) {
            if (x < dim) std::memcpy(&pts[(3 * q + x) * es], &nodes[(dim * q + x) * es], es);
            else std::memset(&pts[(3 * q + x) * es], 0, es);
        }
    std::vector<int64_t> conn(ne * per), offs(ne);
    std::vector<uint8_t> types(ne, (uint8_t)(dim == 2 ? 3 : 5));  // VTK_LINE, VTK_TRIANGLE
    for (int64_t e = 0; e < ne; ++e) {
        for (int v = 0; v < per; ++v) conn[e * per + v] = elems[e * per + v] - 1;
        offs[e] = (e + 1) * per;
    }
    ok = ok && fwrite(&pbytes, 8, 1, f) == 1 && fwrite(pts.data(), 1, pbytes, f) == pbytes;
    ok = ok && fwrite(&cbytes, 8, 1, f) == 1 && fwrite(conn.data(), 1, cbytes, f) == cbytes;
    ok = ok && fwrite(&obytes, 8, 1, f) == 1 && fwrite(offs.data(), 1, obytes, f) == obytes;
    ok = ok && fwrite(&tbytes, 8, 1, f) == 1 && fwrite(types.data(), 1, tbytes, f) == tbytes;
    const char *tail = "\n  </AppendedData>\n</VTKFile>\n";
    ok = ok && fwrite(tail, 1, strlen(tail), f) == strlen(tail);
    fclose(f);
    return ok ? CC_OK : io_fail(CC_ERR_ARGUMENT, "write failed");
}

}  // namespace

extern "C" {

int cc_write_cache_vtk(cc_cache *c, const cc_grid *grid, const char *prefix, int surface) {
    if (!c || !grid || !prefix) return io_fail(CC_ERR_ARGUMENT, "NULL argument");
    Body b;
    int rc = body_of(c, b);
    if (rc) return rc;
    if (b.leaves) return io_fail(CC_ERR_UNSUPPORTED, "a quadtree cache has no image-data layout");
    if (grid->dim != b.dim) return io_fail(CC_ERR_DIMENSION_MISMATCH, "the grid's dimension is not the cache's");
    for (int a = 0; a < b.dim; ++a)
        if (grid->n[a] != b.n[a]) return io_fail(CC_ERR_DIMENSION_MISMATCH, "the grid's n is not the cache's");
    void *sp = nullptr;
    rc = cc_stream(c, &sp);
    if (rc) return rc;
    if (cc_synchronize(c) != CC_OK) return CC_ERR_HIP;
    hipStream_t st = (hipStream_t)sp;
    const int D = b.dim;
    const std::string ft = b.es == 4 ? "Float32" : "Float64";
    std::vector<Array> arr;
    arr.push_back({"volume_fraction", ft, 1, CC_FIELD_VOLUME_FRACTION, 1, 0, b.es, b.es, nullptr, b.ncell, false});
    arr.push_back({"kind", "Int32", 1, CC_FIELD_KIND, 1, 0, 1, 4, i8_to_i32, b.ncell, false});
    arr.push_back({"ambiguous", "Int32", 1, CC_FIELD_AMBIGUOUS, 1, 0, 1, 4, u8_to_i32, b.ncell, false});
    const char *dn[6] = {"-x", "+x", "-y", "+y", "-z", "+z"};
    for (int k = 0; k < 2 * D; ++k)
        arr.push_back({std::string("face_fraction_") + dn[k], ft, 1, CC_FIELD_FACE_FRACTION, 2 * D, k, b.es, b.es,
                       nullptr, b.ncell, false});
    if (b.method == CC_PLIC) {
        arr.push_back({"normal", ft, D, CC_FIELD_NORMALS, D, 0, b.es, b.es, nullptr, b.ncell, false});
        arr.push_back({"intercept", ft, 1, CC_FIELD_INTERCEPTS, 1, 0, b.es, b.es, nullptr, b.ncell, false});
        arr.push_back({"is_valid", "Int32", 1, CC_FIELD_IS_VALID, 1, 0, 1, 4, u8_to_i32, b.ncell, false});
    } else {
        arr.push_back({"centroid", ft, D, CC_FIELD_CENTROID, D, 0, b.es, b.es, nullptr, b.ncell, false});
        arr.push_back({"interface_centroid", ft, D, CC_FIELD_INTERFACE_CENTROID, D, 0, b.es, b.es, nullptr, b.ncell,
                       false});
        arr.push_back({"phi", ft, 1, CC_FIELD_PHI, 1, 0, b.es, b.es, nullptr, b.nnode, true});
    }
    const std::string cells = std::string(prefix) + "_cells.vti";
    rc = write_vti(cells.c_str(), b, grid, arr, st);
    if (rc || !surface) return rc;
    int64_t nn = 0, ne = 0;
    if (b.method == CC_MARCHING_SQUARES) rc = cc_generate_mesh_ms(c, grid, nullptr, &nn, nullptr, &ne);
    else if (b.method == CC_MARCHING_CUBES) rc = cc_generate_mesh_mc(c, grid, nullptr, &nn, nullptr, &ne);
    else rc = cc_generate_mesh_plic(c, grid, 1e-4, nullptr, &nn, nullptr, &ne);
    if (rc) return rc;
    const size_t nes = b.method == CC_MARCHING_SQUARES ? 8 : b.es;  // the contour is Float64 nodes
    std::vector<char> nodes((size_t)nn * D * nes + 1);
    std::vector<int64_t> elems((size_t)ne * D + 1);
    if (b.method == CC_MARCHING_SQUARES)
        rc = cc_generate_mesh_ms(c, grid, (double *)nodes.data(), &nn, elems.data(), &ne);
    else if (b.method == CC_MARCHING_CUBES)
        rc = cc_generate_mesh_mc(c, grid, nodes.data(), &nn, elems.data(), &ne);
    else
        rc = cc_generate_mesh_plic(c, grid, 1e-4, nodes.data(), &nn, elems.data(), &ne);
    if (rc) return rc;
    const std::string surf = std::string(prefix) + "_surface.vtu";
    return write_vtu(surf.c_str(), D, nes == 4 ? CC_FLOAT32 : CC_FLOAT64, nodes, nn, elems, ne);
}

int cc_write_npy(cc_cache *c, int field, const char *path) {
    if (!c || !path) return io_fail(CC_ERR_ARGUMENT, "NULL argument");
    if (field < 0 || field >= CC_NUM_FIELDS) return io_fail(CC_ERR_ARGUMENT, "unknown field");
    Body b;
    int rc = body_of(c, b);
    if (rc) return rc;
    void *dptr = nullptr;
    int64_t nbytes = 0;
    rc = cc_column(c, field, &dptr, &nbytes);
    if (rc) return rc;
    if (!dptr || nbytes == 0) return io_fail(CC_ERR_UNSUPPORTED, "this cache has no such column");
    const size_t es = (field == CC_FIELD_KIND || field == CC_FIELD_AMBIGUOUS || field == CC_FIELD_IS_VALID) ? 1 : b.es;
    const char *descr = field == CC_FIELD_KIND ? "|i1" : (es == 1 ? "|u1" : (es == 4 ? "<f4" : "<f8"));
    const int D = b.dim;
    // components first (an SVector of SVectors: face_centroid_local is (D-1, 2D, cells...)), then cells
    std::vector<int64_t> dims;
    if (field == CC_FIELD_FACE_CENTROID_LOCAL) {
        dims.push_back(D - 1);
        dims.push_back(2 * D);
    } else if (field == CC_FIELD_FACE_FRACTION) {
        dims.push_back(2 * D);
    } else if (field == CC_FIELD_CENTROID || field == CC_FIELD_INTERFACE_CENTROID || field == CC_FIELD_NORMALS) {
        dims.push_back(D);
    } else if (field == CC_FIELD_PHI && b.leaves) {
        dims.push_back(4);  // four corner values per leaf
    }
    if (b.leaves) {
        dims.push_back(b.ncell);
    } else {
        for (int a = 0; a < D; ++a) dims.push_back(field == CC_FIELD_PHI ? b.n[a] + 1 : b.n[a]);
    }
    std::string shape = "(";
    for (size_t q = 0; q < dims.size(); ++q) shape += std::to_string(dims[q]) + (dims.size() == 1 ? ",)" : (q + 1 < dims.size() ? ", " : ")"));
    std::string hdr = std::string("{'descr': '") + descr + "', 'fortran_order': True, 'shape': " + shape + ", }";
    while ((10 + hdr.size() + 1) % 64) hdr += ' ';
    hdr += '\n';
    FILE *f = fopen(path, "wb");
    if (!f) return io_fail(CC_ERR_ARGUMENT, std::string("cannot open ") + path);
    const unsigned char magic[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
    const uint16_t hl = (uint16_t)hdr.size();
    bool ok = fwrite(magic, 1, 8, f) == 8 && fwrite(&hl, 2, 1, f) == 1 && fwrite(hdr.data(), 1, hdr.size(), f) == hdr.size();
    void *sp = nullptr;
    rc = ok ? cc_stream(c, &sp) : io_fail(CC_ERR_ARGUMENT, "write failed");
    if (rc == CC_OK && cc_synchronize(c) != CC_OK) rc = CC_ERR_HIP;
    if (rc == CC_OK) {
        const int64_t count = nbytes / (int64_t)es, chunk = count < CHUNK ? (count > 0 ? count : 1) : CHUNK;
        std::vector<char> host((size_t)chunk * es);
        rc = stream_column(f, dptr, count, 1, 0, es, (hipStream_t)sp, nullptr, host, chunk);
    }
    fclose(f);
    return rc;
}

}  // extern "C"
