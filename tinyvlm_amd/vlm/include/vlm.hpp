// vlm.hpp -- the steady caller of the hot path: flow data, plot3d mesh IO and the VLM driver
// (reference: vlm/include/vlm_data.hpp:10-37, vlm/include/vlm_mesh.hpp:14-20,
// vlm/include/vlm.hpp:41-97).  The driver is templated on the working precision; VLM is the
// reference's FP32 solver, VLMd the same sequence in FP64.
#pragma once

#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "linalg3.hpp"
#include "vlm_backend.hpp"

namespace vlm {

// freestream = u_inf (cos a cos b, -cos a sin b, sin a); lift axis = normalize(fs x y); stream
// axis = normalize(fs) (vlm/src/vlm_utils.cpp:5-19)
linalg::float3 compute_freestream(f32 u_inf, f32 alpha, f32 beta);
linalg::float3 compute_lift_axis(const linalg::float3& freestream);
linalg::float3 compute_stream_axis(const linalg::float3& freestream);
constexpr f32 to_degrees(f32 radians) { return radians * 180.0f / PI_f; }
constexpr f32 to_radians(f32 degrees) { return degrees * PI_f / 180.0f; }

class FlowData {
   public:
    f32 alpha, beta, u_inf, rho;
    linalg::float3 freestream, lift_axis, stream_axis;
    FlowData(f32 alpha_, f32 beta_, f32 u_inf_, f32 rho_);
    FlowData(const linalg::float3& freestream_, f32 rho_);
};

class AeroCoefficients {
   public:
    f32 cl, cd;
    linalg::float3 cm;
    AeroCoefficients(f32 cl_, f32 cd_, const linalg::float3& cm_) : cl(cl_), cd(cd_), cm(cm_) {}
};

using SurfDims = std::pair<i64, i64>;  // (nc, ns) panels

// single-block 2-D plot3d: "1", "ni nj 1", then all x, all y, all z with i (chordwise) fastest;
// vertices{ns+1, nc+1, 4}(j, i, k).  The quarter-chord shift moves every chordwise row a quarter
// panel aft and extrapolates the trailing edge (vlm/src/vlm_mesh.cpp:15-32, :34-91).
class MeshIO {
   public:
    explicit MeshIO(const std::string& format);
    SurfDims get_dims(const std::string& filename) const;
    void read(const std::string& filename, const TensorView3fH& vertices, bool qc = true) const;
    void read(const std::string& filename, const TensorView3dH& vertices, bool qc = true) const;
};

template <typename T>
class SimulationT {
   public:
    std::unique_ptr<Backend> backend;
    MultiTensor<T, 3, Location::Device> verts_wing_init{backend->memory.get()};
    MultiTensor<T, 3, Location::Host> verts_wing_init_h{backend->memory.get()};
    MultiTensor<T, 3, Location::Device> verts_wing{backend->memory.get()};
    MultiTensor<T, 3, Location::Device> verts_wake{backend->memory.get()};
    MultiDim<2> assembly_wings;  // (ns, nc) per surface

    // meshes are parsed and quarter-chord shifted in FP32 (the reference drivers' precision), then
    // widened when T is double (vlm/src/vlm.cpp:17-44)
    SimulationT(const std::string& backend_name, const std::vector<std::string>& meshes, bool qc = true);
    virtual ~SimulationT() = default;
};

template <typename T>
class VLMT final : public SimulationT<T> {
    using S = SimulationT<T>;

   public:
    VLMT(const std::string& backend_name, const std::vector<std::string>& meshes);
    // one steady solve (vlm/src/vlm.cpp:97-151).  CL as the reference; CM about ref_pt (in the
    // wing frame, moved with the wing) is computed too, the reference returns zeros there.
    AeroCoefficients run(const FlowData& flow, linalg::vec3<T> ref_pt = {0.25, 0.0, 0.0});
    // the last run's coefficients in the working precision (AeroCoefficients holds FP32)
    T last_cl = 0;
    linalg::vec3<T> last_cm{};

    std::vector<i32> condition0;
    MultiTensor<T, 3, Location::Device> colloc_d{S::backend->memory.get()};
    MultiTensor<T, 3, Location::Device> normals_d{S::backend->memory.get()};
    MultiTensor<T, 2, Location::Device> areas_d{S::backend->memory.get()};
    Tensor<T, 2, Location::Device> lhs{S::backend->memory.get()};
    Tensor<T, 1, Location::Device> rhs{S::backend->memory.get()};
    MultiTensor<T, 2, Location::Device> gamma_wing{S::backend->memory.get()};
    MultiTensor<T, 2, Location::Device> gamma_wake{S::backend->memory.get()};
    MultiTensor<T, 2, Location::Device> gamma_wing_prev{S::backend->memory.get()};
    MultiTensor<T, 2, Location::Device> gamma_wing_delta{S::backend->memory.get()};
    MultiTensor<T, 3, Location::Device> forces{S::backend->memory.get()};
    MultiTensor<T, 3, Location::Device> local_velocities{S::backend->memory.get()};
    Tensor<T, 2, Location::Host> wake_transform{S::backend->memory.get()};
    MultiTensor<T, 2, Location::Device> transforms{S::backend->memory.get()};

   private:
    std::unique_ptr<LU> solver;
    void alloc_buffers();
};

using VLM = VLMT<f32>;
using VLMd = VLMT<f64>;

}  // namespace vlm
