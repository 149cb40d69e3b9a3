// rsd_io.hip -- the reference's binary mesh (PMSH v1) and nodal field
// (PNDF v1) files, host side.
//
// PMSH (src/io/MeshIO.cpp:31-140, io/MeshIO.hpp:49-50), little endian:
//   u32 magic 0x504D5348, u32 version 1,
//   u32 parametric dim, spatial dim, element family, basis type, nodes,
//       elements, nodes per element, faces per element, number of orders,
//   u32 basis order x number of orders,
//   u8 has extraction operators, u64 operator size,
//   f64 xyz [nodes * spatial dim], u64 ien [elements * nodes per element],
//   i32 rng [elements * faces per element], f64 operators [size].
// PNDF (io/fieldio/FieldIO.hpp:19-26, FieldIO.tpp:173-231):
//   u32 magic 0x504E4446, u32 version 1, u64 nodes, u32 dofs per node,
//   f64 time, u64 mesh tag, f64 values [nodes * dofs];
//   mesh tag: FNV-1a over the bytes of xyz (f64) and then ien (u64).
#include "rsd_problem.hpp"

#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

namespace rsd {

constexpr uint32_t kPmshMagic = 0x504D5348u, kPmshVersion = 1u;
constexpr uint32_t kPndfMagic = 0x504E4446u, kPndfVersion = 1u;

// little-endian scalar I/O on a stdio file (x86-64 and gfx hosts are LE)
struct File {
    FILE* f = nullptr;
    std::string path;
    File(const char* p, const char* mode) : path(p ? p : "") {
        RSD_CHECK(p, "null path");
        f = std::fopen(p, mode);
        RSD_CHECK(f, "cannot open file '" + path + "'");
    }
    ~File() {
        if (f) std::fclose(f);
    }
    template<typename T>
    T get() {
        T v;
        RSD_CHECK(std::fread(&v, sizeof(T), 1, f) == 1, "unexpected end of file in '" + path + "'");
        return v;
    }
    template<typename T>
    void get_n(T* p, size_t n) {
        RSD_CHECK(std::fread(p, sizeof(T), n, f) == n, "unexpected end of file in '" + path + "'");
    }
    template<typename T>
    void put(T v) {
        RSD_CHECK(std::fwrite(&v, sizeof(T), 1, f) == 1, "write failed on '" + path + "'");
    }
    template<typename T>
    void put_n(const T* p, size_t n) {
        RSD_CHECK(std::fwrite(p, sizeof(T), n, f) == n, "write failed on '" + path + "'");
    }
};

struct PmshHeader {
    uint32_t pdim, sdim, family, basis, nn, ne, npe, faces, n_orders;
    uint32_t orders[8];
    uint8_t has_c;
    uint64_t c_size;
};

static PmshHeader read_pmsh_header(File& F) {
    RSD_CHECK(F.get<uint32_t>() == kPmshMagic, "bad magic number: '" + F.path + "' is not a PMSH file");
    const uint32_t ver = F.get<uint32_t>();
    RSD_CHECK(ver == kPmshVersion, "unsupported PMSH version " + std::to_string(ver));
    PmshHeader h{};
    h.pdim = F.get<uint32_t>();
    h.sdim = F.get<uint32_t>();
    h.family = F.get<uint32_t>();
    h.basis = F.get<uint32_t>();
    h.nn = F.get<uint32_t>();
    h.ne = F.get<uint32_t>();
    h.npe = F.get<uint32_t>();
    h.faces = F.get<uint32_t>();
    h.n_orders = F.get<uint32_t>();
    RSD_CHECK(h.n_orders <= 8, "PMSH: more than 8 basis orders");
    uint64_t npe = 1;
    for (uint32_t i = 0; i < h.n_orders; ++i) {
        h.orders[i] = F.get<uint32_t>();
        RSD_CHECK(h.orders[i] > 0, "PMSH: zero basis order");
        npe *= h.orders[i] + 1;
    }
    h.has_c = F.get<uint8_t>();
    h.c_size = F.get<uint64_t>();
    // Mesh::isValid (mesh/Mesh.hpp:41-65)
    RSD_CHECK(h.pdim > 0 && h.sdim > 0, "PMSH: zero dimension");
    RSD_CHECK(npe == h.npe, "PMSH: nodes per element do not match the basis orders");
    return h;
}

// FieldIO::computeMeshTag (FieldIO.tpp:173-190)
static uint64_t mesh_tag(const std::vector<double>& xyz, const std::vector<int32_t>& ien) {
    uint64_t h = 14695981039346656037ull;
    auto mix = [&h](const void* data, size_t bytes) {
        const uint8_t* p = static_cast<const uint8_t*>(data);
        for (size_t i = 0; i < bytes; ++i) {
            h ^= p[i];
            h *= 1099511628211ull;
        }
    };
    mix(xyz.data(), xyz.size() * sizeof(double));
    for (int32_t v : ien) {
        const uint64_t w = (uint64_t)(int64_t)v;
        mix(&w, sizeof(w));
    }
    return h;
}

}  // namespace rsd

using namespace rsd;

extern "C" {

int rsd_pmsh_info(const char* path, int64_t* out) {
    return guard([&] {
        RSD_CHECK(out, "rsd_pmsh_info: null out");
        File F(path, "rb");
        const PmshHeader h = read_pmsh_header(F);
        const int64_t v[9] = {h.pdim, h.sdim, h.family, h.basis, h.nn, h.ne, h.npe, h.faces, h.n_orders};
        for (int i = 0; i < 9; ++i) out[i] = v[i];
        for (int i = 0; i < 3; ++i) out[9 + i] = i < (int)h.n_orders ? h.orders[i] : 0;
    });
}

int rsd_pmsh_read(const char* path, double* xyz, int64_t* ien, int32_t* rng) {
    return guard([&] {
        RSD_CHECK(xyz && ien && rng, "rsd_pmsh_read: null argument");
        File F(path, "rb");
        const PmshHeader h = read_pmsh_header(F);
        F.get_n(xyz, (size_t)h.nn * h.sdim);
        std::vector<uint64_t> tmp((size_t)h.ne * h.npe);
        F.get_n(tmp.data(), tmp.size());
        for (size_t k = 0; k < tmp.size(); ++k) {
            RSD_CHECK(tmp[k] < h.nn, "PMSH: connectivity entry out of range");
            ien[k] = (int64_t)tmp[k];
        }
        F.get_n(rng, (size_t)h.ne * h.faces);
    });
}

int rsd_pmsh_write(const char* path, int32_t pdim, int32_t sdim, int32_t family, int32_t basis_type,
                   int64_t n_nodes, int64_t n_elem, int32_t npe, int32_t faces, int32_t n_orders,
                   const int32_t* orders, const double* xyz, const int64_t* ien, const int32_t* rng) {
    return guard([&] {
        RSD_CHECK(xyz && ien && rng && (n_orders == 0 || orders), "rsd_pmsh_write: null argument");
        RSD_CHECK(pdim > 0 && sdim > 0 && n_nodes >= 0 && n_elem >= 0 && n_orders >= 0, "rsd_pmsh_write: bad sizes");
        RSD_CHECK(n_nodes <= UINT32_MAX && n_elem <= UINT32_MAX, "rsd_pmsh_write: PMSH v1 counts are 32-bit");
        int64_t np = 1;
        for (int i = 0; i < n_orders; ++i) {
            RSD_CHECK(orders[i] > 0, "rsd_pmsh_write: zero basis order");
            np *= orders[i] + 1;
        }
        RSD_CHECK(np == npe, "rsd_pmsh_write: nodes per element do not match the basis orders");
        File F(path, "wb");
        F.put<uint32_t>(kPmshMagic);
        F.put<uint32_t>(kPmshVersion);
        const uint32_t hdr[9] = {(uint32_t)pdim, (uint32_t)sdim, (uint32_t)family, (uint32_t)basis_type,
                                 (uint32_t)n_nodes, (uint32_t)n_elem, (uint32_t)npe, (uint32_t)faces,
                                 (uint32_t)n_orders};
        F.put_n(hdr, 9);
        for (int i = 0; i < n_orders; ++i) F.put<uint32_t>((uint32_t)orders[i]);
        F.put<uint8_t>(0);   // no extraction operators (Lagrange meshes)
        F.put<uint64_t>(0);
        F.put_n(xyz, (size_t)n_nodes * sdim);
        std::vector<uint64_t> tmp((size_t)n_elem * npe);
        for (size_t k = 0; k < tmp.size(); ++k) tmp[k] = (uint64_t)ien[k];
        F.put_n(tmp.data(), tmp.size());
        F.put_n(rng, (size_t)n_elem * faces);
    });
}

int rsd_problem_create_pmsh(rsd_ctx* ctx, const char* path, int32_t nq, int32_t n_bc, const int32_t* bc_tags,
                            const char* const* bc_exprs, rsd_problem** out) {
    return guard([&] {
        RSD_CHECK(ctx && out, "rsd_problem_create_pmsh: null argument");
        PmshHeader h;
        {
            File F(path, "rb");
            h = read_pmsh_header(F);
        }
        // ElementFamily Quad = 0, Hex = 2; BasisType Lagrange = 0 (mesh/ElementFamily.hpp, BasisType.hpp)
        RSD_CHECK(h.basis == 0 && !h.has_c, "PMSH: only Lagrange meshes run on the device");
        RSD_CHECK((h.family == 0 && h.pdim == 2 && h.sdim == 2 && h.n_orders == 2) ||
                      (h.family == 2 && h.pdim == 3 && h.sdim == 3 && h.n_orders == 3),
                  "PMSH: quadrilaterals in 2D or hexahedra in 3D");
        std::vector<double> xyz((size_t)h.nn * h.sdim);
        std::vector<int64_t> ien((size_t)h.ne * h.npe);
        std::vector<int32_t> rng((size_t)h.ne * h.faces);
        RSD_CHECK(rsd_pmsh_read(path, xyz.data(), ien.data(), rng.data()) == RSD_OK, rsd_last_error_string());
        int32_t p[3] = {1, 1, 1};
        for (uint32_t i = 0; i < h.n_orders; ++i) p[i] = (int32_t)h.orders[i];
        RSD_CHECK(rsd_problem_create(ctx, (int32_t)h.pdim, p, nq, h.nn, h.ne, xyz.data(), ien.data(), rng.data(),
                                     n_bc, bc_tags, bc_exprs, out) == RSD_OK,
                  rsd_last_error_string());
    });
}

int rsd_problem_nodal_field(rsd_ctx* ctx, rsd_problem* pb, double time, const double* d_U, double* nodal) {
    return guard([&] {
        RSD_CHECK(ctx && pb && d_U && nodal, "rsd_problem_nodal_field: null argument");
        RSD_CHECK(ctx == pb->ctx, "problem belongs to another context");
        // FieldIO::reconstructNodalField (FieldIO.tpp:126-171): free values
        // from U, constrained ones from the Dirichlet values at `time`
        std::vector<double> U((size_t)pb->n_free), g((size_t)pb->n_nodes);
        update_gval(pb, time);
        if (pb->n_free)
            RSD_HIP(hipMemcpyAsync(U.data(), d_U, sizeof(double) * U.size(), hipMemcpyDeviceToHost, ctx->stream));
        RSD_HIP(hipMemcpyAsync(g.data(), pb->gval.p, sizeof(double) * g.size(), hipMemcpyDeviceToHost, ctx->stream));
        RSD_HIP(hipStreamSynchronize(ctx->stream));
        for (int64_t v = 0; v < pb->n_nodes; ++v) nodal[v] = pb->h_t2a[v] < 0 ? g[v] : 0.0;
        for (int64_t i = 0; i < pb->n_free; ++i) nodal[pb->h_a2t[i]] = U[i];
    });
}

int rsd_pndf_write(rsd_ctx* ctx, rsd_problem* pb, double time, const double* d_U, const char* path) {
    return guard([&] {
        RSD_CHECK(ctx && pb && d_U && path, "rsd_pndf_write: null argument");
        std::vector<double> nodal((size_t)pb->n_nodes);
        RSD_CHECK(rsd_problem_nodal_field(ctx, pb, time, d_U, nodal.data()) == RSD_OK, rsd_last_error_string());
        File F(path, "wb");
        F.put<uint32_t>(kPndfMagic);
        F.put<uint32_t>(kPndfVersion);
        F.put<uint64_t>((uint64_t)pb->n_nodes);
        F.put<uint32_t>(1u);
        F.put<double>(time);
        F.put<uint64_t>(mesh_tag(pb->h_xyz, pb->h_ien));
        F.put_n(nodal.data(), nodal.size());
    });
}

int rsd_pndf_read(rsd_problem* pb, const char* path, double* nodal, double* time) {
    return guard([&] {
        RSD_CHECK(pb && nodal, "rsd_pndf_read: null argument");
        File F(path, "rb");
        RSD_CHECK(F.get<uint32_t>() == kPndfMagic, "bad magic number: '" + F.path + "' is not a PNDF file");
        const uint32_t ver = F.get<uint32_t>();
        RSD_CHECK(ver == kPndfVersion, "unsupported PNDF version " + std::to_string(ver));
        const uint64_t nn = F.get<uint64_t>();
        const uint32_t dofs = F.get<uint32_t>();
        const double t = F.get<double>();
        const uint64_t tag = F.get<uint64_t>();
        RSD_CHECK(nn == (uint64_t)pb->n_nodes && dofs == 1u, "PNDF: node or DOF count does not match the problem");
        RSD_CHECK(tag == mesh_tag(pb->h_xyz, pb->h_ien), "PNDF: mesh tag does not match the problem's mesh");
        F.get_n(nodal, (size_t)nn);
        if (time) *time = t;
    });
}

int rsd_problem_mesh_tag(rsd_problem* pb, uint64_t* out) {
    return guard([&] {
        RSD_CHECK(pb && out, "rsd_problem_mesh_tag: null argument");
        *out = mesh_tag(pb->h_xyz, pb->h_ien);
    });
}

}  // extern "C"
