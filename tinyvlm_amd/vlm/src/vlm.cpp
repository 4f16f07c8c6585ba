// vlm.cpp -- flow data, plot3d mesh IO and the steady VLM driver over the Backend interface
// (reference: vlm/src/vlm_data.cpp:8-23, vlm/src/vlm_utils.cpp:5-19, vlm/src/vlm_mesh.cpp:15-130,
// vlm/src/vlm.cpp:17-151).  Every array operation is a Backend / BLAS / LU call; the host only
// parses meshes and builds the 4x4 wake transform.
#include "vlm.hpp"

#include <cmath>
#include <cstdio>
#include <fstream>
#include <stdexcept>

namespace vlm {

linalg::float3 compute_freestream(f32 u_inf, f32 alpha, f32 beta) {
    return {u_inf * std::cos(alpha) * std::cos(beta), -u_inf * std::cos(alpha) * std::sin(beta), u_inf * std::sin(alpha)};
}
linalg::float3 compute_lift_axis(const linalg::float3& fs) { return linalg::normalize(linalg::cross(fs, {0.f, 1.f, 0.f})); }
linalg::float3 compute_stream_axis(const linalg::float3& fs) { return linalg::normalize(fs); }

FlowData::FlowData(f32 alpha_, f32 beta_, f32 u_inf_, f32 rho_)
    : alpha(alpha_), beta(beta_), u_inf(u_inf_), rho(rho_), freestream(compute_freestream(u_inf_, alpha_, beta_)),
      lift_axis(compute_lift_axis(freestream)), stream_axis(compute_stream_axis(freestream)) {}

FlowData::FlowData(const linalg::float3& fs, f32 rho_)
    : alpha(std::atan(fs.z / std::sqrt(fs.x * fs.x + fs.y * fs.y))), beta(std::atan(fs.y / fs.x)), u_inf(linalg::length(fs)),
      rho(rho_), freestream(fs), lift_axis(compute_lift_axis(fs)), stream_axis(compute_stream_axis(fs)) {}

// ---- plot3d ----------------------------------------------------------------------------------

namespace {

SurfDims plot3d_dims(std::ifstream& f) {
    i64 blocks = 0, ni = 0, nj = 0, nk = 0;
    f >> blocks;
    if (blocks != 1) throw std::runtime_error("Only single block plot3d mesh is supported");
    f >> ni >> nj >> nk;
    if (nk != 1) throw std::runtime_error("Only 2D plot3d mesh is supported");
    if (!f || ni < 2 || nj < 2) throw std::runtime_error("Malformed plot3d header");
    return {ni - 1, nj - 1};
}

template <typename T>
void plot3d_read(std::ifstream& f, const TensorView<T, 3, Location::Host>& v) {
    const auto [nc, ns] = plot3d_dims(f);
    if (v.shape(0) != ns + 1 || v.shape(1) != nc + 1 || v.shape(2) != 4)
        throw std::invalid_argument("plot3d: vertex tensor does not match the mesh dimensions");
    for (i64 k = 0; k < 3; k++)  // x block, y block, z block; chordwise index fastest
        for (i64 j = 0; j <= ns; j++)
            for (i64 i = 0; i <= nc; i++) {
                T x;
                if (!(f >> x)) throw std::runtime_error("plot3d: truncated coordinate block");
                v(j, i, k) = x;
            }
    for (i64 i = 1; i <= nc; i++)
        if (v(0, i, 1) != v(0, 0, 1)) throw std::runtime_error("Mesh vertices should be ordered in chordwise direction");
}

// vortex rings at the quarter chord: every interior row moves a quarter panel aft, the trailing
// edge is extrapolated so the last ring ends a quarter panel behind it
template <typename T>
void quarterchord(const TensorView<T, 3, Location::Host>& v) {
    const i64 nv = v.shape(1);
    for (i64 i = 0; i + 1 < nv; i++)
        for (i64 j = 0; j < v.shape(0); j++)
            for (i64 k = 0; k < 3; k++) v(j, i, k) = 0.75f * v(j, i, k) + 0.25f * v(j, i + 1, k);
    for (i64 j = 0; j < v.shape(0); j++)
        for (i64 k = 0; k < 3; k++) v(j, nv - 1, k) = (4.f / 3.f) * v(j, nv - 1, k) - (1.f / 3.f) * v(j, nv - 2, k);
}

template <typename T>
void read_impl(const std::string& filename, const TensorView<T, 3, Location::Host>& v, bool qc) {
    std::ifstream f(filename);
    if (!f.is_open()) throw std::runtime_error("Failed to open mesh file: " + filename);
    plot3d_read(f, v);
    if (qc) quarterchord(v);
}

}  // namespace

MeshIO::MeshIO(const std::string& format) {
    if (format != "plot3d") throw std::invalid_argument("MeshIO: only the plot3d format is supported");
}

SurfDims MeshIO::get_dims(const std::string& filename) const {
    std::ifstream f(filename);
    if (!f.is_open()) throw std::runtime_error("Failed to open mesh file: " + filename);
    return plot3d_dims(f);
}

void MeshIO::read(const std::string& filename, const TensorView3fH& v, bool qc) const { read_impl(filename, v, qc); }
void MeshIO::read(const std::string& filename, const TensorView3dH& v, bool qc) const { read_impl(filename, v, qc); }

// ---- simulation ------------------------------------------------------------------------------

template <typename T>
SimulationT<T>::SimulationT(const std::string& backend_name, const std::vector<std::string>& meshes, bool qc)
    : backend(create_backend(backend_name)) {
    const MeshIO io{"plot3d"};
    MultiDim<3> shapes;
    for (const auto& name : meshes) {
        const auto [nc, ns] = io.get_dims(name);
        assembly_wings.push_back({ns, nc});
        shapes.push_back({ns + 1, nc + 1, 4});
    }
    verts_wing_init.init(shapes);
    verts_wing_init_h.init(shapes);
    verts_wing.init(shapes);
    for (size_t m = 0; m < meshes.size(); m++) {
        Tensor<f32, 3, Location::Host> parsed{backend->memory.get()};
        parsed.init(shapes[m]);
        io.read(meshes[m], parsed.view(), qc);
        const auto& h = verts_wing_init_h.views()[m];
        for (i64 j = 0; j < h.shape(0); j++)
            for (i64 i = 0; i < h.shape(1); i++) {
                for (i64 k = 0; k < 3; k++) h(j, i, k) = (T)parsed.view()(j, i, k);
                h(j, i, 3) = 1;
            }
        h.to(verts_wing_init.views()[m]);
    }
}

template <typename T>
VLMT<T>::VLMT(const std::string& backend_name, const std::vector<std::string>& meshes) : SimulationT<T>(backend_name, meshes) {
    solver = S::backend->create_lu_solver();
    alloc_buffers();
}

// buffers of vlm/src/vlm.cpp:60-95: one row of 100-chord wake panels per surface
template <typename T>
void VLMT<T>::alloc_buffers() {
    i64 n = 0;
    MultiDim<3> p3, wake3;
    MultiDim<2> p2, wake2, tr2;
    for (const auto& [ns, nc] : S::assembly_wings) {
        n += ns * nc;
        p3.push_back({ns, nc, 3});
        p2.push_back({ns, nc});
        wake2.push_back({ns, 1});
        wake3.push_back({ns + 1, 2, 4});
        tr2.push_back({4, 4});
    }
    condition0.resize(S::assembly_wings.size() * S::assembly_wings.size());
    normals_d.init(p3);
    colloc_d.init(p3);
    areas_d.init(p2);
    S::verts_wake.init(wake3);
    lhs.init({n, n});
    rhs.init({n});
    gamma_wing.init(p2);
    gamma_wing_prev.init(p2);
    gamma_wing_delta.init(p2);
    gamma_wake.init(wake2);
    local_velocities.init(p3);
    forces.init(p3);
    wake_transform.init({4, 4});
    transforms.init(tr2);
    solver->init(lhs.view());
    for (const auto& w : S::verts_wake.views()) w.slice(All, All, 3).fill((T)1);
}

template <typename T>
AeroCoefficients VLMT<T>::run(const FlowData& flow, linalg::vec3<T> ref_pt) {
    Backend& b = *S::backend;
    lhs.view().fill(0);
    rhs.view().fill(0);
    for (size_t m = 0; m < S::verts_wing.views().size(); m++) S::verts_wing_init.views()[m].to(S::verts_wing.views()[m]);

    // move the wing 100 chords along -freestream so that the shed row is a long wake panel
    const T alpha = flow.alpha, beta = flow.beta, u = flow.u_inf;
    const auto pose = linalg::mat4<T>::translation(-100 * u * std::cos(alpha), 0, -100 * u * std::sin(alpha));
    const auto& wt = wake_transform.view();
    for (i64 c = 0; c < 4; c++)
        for (i64 r = 0; r < 4; r++) wt(r, c) = pose.m[c * 4 + r];
    for (const auto& t : transforms.views()) wt.to(t);

    b.wake_shed(S::verts_wing.views(), S::verts_wake.views(), 0);
    b.displace_wing(transforms.views(), S::verts_wing.views(), S::verts_wing_init.views());
    b.wake_shed(S::verts_wing.views(), S::verts_wake.views(), 1);
    b.mesh_metrics(alpha, S::verts_wing.views(), colloc_d.views(), normals_d.views(), areas_d.views());
    b.lhs_assemble(lhs.view(), colloc_d.views(), normals_d.views(), S::verts_wing.views(), S::verts_wake.views(), condition0, 1);

    linalg::vec3<T> fs{u * std::cos(alpha) * std::cos(beta), -u * std::cos(alpha) * std::sin(beta), u * std::sin(alpha)};
    for (const auto& vel : local_velocities.views()) {
        vel.slice(All, All, 0).fill(fs.x);
        vel.slice(All, All, 1).fill(fs.y);
        vel.slice(All, All, 2).fill(fs.z);
    }
    b.rhs_assemble_velocities(rhs.view(), normals_d.views(), local_velocities.views());

    solver->factorize(lhs.view());
    solver->solve(lhs.view(), rhs.view().reshape(rhs.size(), 1));

    // circulation per surface, its chordwise difference (ring strength seen by each bound
    // segment) and the trailing-edge value shed to the wake (vlm/src/vlm.cpp:126-139)
    i64 begin = 0;
    for (size_t m = 0; m < S::assembly_wings.size(); m++) {
        const auto& g = gamma_wing.views()[m];
        const auto& gd = gamma_wing_delta.views()[m];
        const i64 end = begin + g.size();
        rhs.view().slice(Range{begin, end}).to(g.reshape(g.size()));
        g.to(gamma_wing_prev.views()[m]);
        g.to(gd);
        g.slice(All, -1).to(gamma_wake.views()[m].slice(All, -1));
        if (g.shape(1) > 1) b.blas->axpy((T)-1, g.slice(All, Range{0, -2}), gd.slice(All, Range{1, -1}));
        begin = end;
    }

    b.forces_steady_multibody(S::verts_wing.views(), gamma_wing_delta.views(), local_velocities.views(), forces.views());
    const T cl = b.coeff_cl_multibody(forces.views(), areas_d.views(), fs, (T)flow.rho);
    auto ref = pose.apply(ref_pt);
    const auto cm = b.coeff_cm_multibody(forces.views(), S::verts_wing.views(), areas_d.views(), ref, fs, (T)flow.rho);
    last_cl = cl;
    last_cm = cm;
    return AeroCoefficients{(f32)cl, 0.0f, {(f32)cm.x, (f32)cm.y, (f32)cm.z}};
}

template class SimulationT<f32>;
template class SimulationT<f64>;
template class VLMT<f32>;
template class VLMT<f64>;

}  // namespace vlm
