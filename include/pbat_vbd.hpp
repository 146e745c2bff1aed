/*
 * pbat_vbd.hpp -- header-only C++ wrapper over the C ABI of pbat_vbd.h.
 *
 * Mirrors the class and method names of the reference's C++ API so calling
 * code ports by changing the namespace:
 *   pbat::sim::vbd::Data        (source/pbat/sim/vbd/Data.h:14-247)
 *   pbat::gpu::vbd::Integrator  (source/pbat/gpu/vbd/Integrator.h:33-150)
 * become pbat_amd::sim::vbd::Data and pbat_amd::gpu::vbd::Integrator.
 *
 * Matrices are column-major with the reference's shapes (3 x nV positions,
 * 4 x nT elements). Any type with rows(), cols() and data() in column-major
 * order (e.g. an Eigen matrix) can be passed; results come back as Matrix,
 * a minimal column-major owner. Errors throw std::invalid_argument or
 * std::runtime_error like the reference.
 */
#ifndef PBAT_VBD_HPP
#define PBAT_VBD_HPP

#include "pbat_vbd.h"

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace pbat_amd {

inline void Check(int status)
{
    if (status == PBAT_VBD_OK)
        return;
    std::string const what = pbat_vbd_last_error();
    if (status == PBAT_VBD_ERROR_INVALID_ARGUMENT || status == PBAT_VBD_ERROR_SINGULAR_ELEMENT)
        throw std::invalid_argument(what);
    throw std::runtime_error(what);
}

/* Column-major rows x cols matrix. */
template <class T>
struct Matrix
{
    int64_t nRows = 0, nCols = 0;
    std::vector<T> values;
    Matrix() = default;
    Matrix(int64_t r, int64_t c, T v = T{}) : nRows(r), nCols(c), values(static_cast<size_t>(r * c), v) {}
    int64_t rows() const { return nRows; }
    int64_t cols() const { return nCols; }
    T* data() { return values.data(); }
    T const* data() const { return values.data(); }
    T& operator()(int64_t r, int64_t c) { return values[static_cast<size_t>(c * nRows + r)]; }
    T operator()(int64_t r, int64_t c) const { return values[static_cast<size_t>(c * nRows + r)]; }
};

template <class T, class TMatrix>
Matrix<T> ToMatrix(TMatrix const& A)
{
    Matrix<T> M(static_cast<int64_t>(A.rows()), static_cast<int64_t>(A.cols()));
    auto const* src = A.data();
    for (size_t k = 0; k < M.values.size(); ++k)
        M.values[k] = static_cast<T>(src[k]);
    return M;
}

namespace sim::vbd {

/* sim/vbd/Enums.h */
enum class EInitializationStrategy { Position, Inertia, KineticEnergyMinimum, AdaptiveVbd, AdaptivePbat };

/* sim/vbd/Data.h: the fields consumed by the GPU integrator, built by Construct(). */
struct Data
{
    Matrix<double> X, x, v, aext, GP, lame;
    Matrix<int64_t> E;
    std::vector<double> m, wg, rhoe;
    std::vector<int64_t> GVGp, GVGe, GVGilocal, dbc, colors, Pptr, Padj;
    int ordering = PBAT_VBD_ORDER_LARGEST_DEGREE;
    int selection = PBAT_VBD_SELECT_LEAST_USED;
    EInitializationStrategy strategy = EInitializationStrategy::AdaptivePbat;
    double kD = 0.;
    double detHZero = 1e-7;
    int32_t acceleration = PBAT_VBD_ACCELERATION_NONE; // sim/vbd/Data.h:231
    double rho = 1.;                                   // sim/vbd/Data.h:234
    int32_t window = 5;                                // sim/vbd/Data.h:237
    double eta = 0.2;                                  // sim/vbd/Data.h:244
    double tau = 2.;                                   // sim/vbd/Data.h:245
    bool curved = true;                                // sim/vbd/Data.h:246
    std::vector<double> mue, lambdae;
    // collision mesh and contact model (sim/vbd/Data.h:165-178, 225-229)
    std::vector<int64_t> B, V;
    Matrix<int64_t> F;
    std::vector<double> XVA, FA;
    double muC = 1e6, muF = 0.3, epsv = 1e-3;
    int32_t activeSetUpdateFrequency = 1;

    template <class TX, class TE>
    Data& WithVolumeMesh(TX const& Xin, TE const& Ein)
    {
        if (Xin.rows() != 3 || Ein.rows() != 4)
            throw std::invalid_argument("expected 3 x |#nodes| positions and 4 x |#elements| elements");
        X = ToMatrix<double>(Xin);
        E = ToMatrix<int64_t>(Ein);
        return *this;
    }
    /** sim/vbd/Data.cpp:34-54: collision vertices V and triangles F (3 x |#triangles|);
     * vertex areas get a sixth of each incident triangle's double area, triangle
     * areas half of it. Call after WithVolumeMesh. */
    template <class TF>
    Data& WithSurfaceMesh(std::vector<int64_t> Vin, TF const& Fin)
    {
        if (Fin.rows() != 3)
            throw std::invalid_argument("expected 3 x |#collision triangles| triangles");
        V = std::move(Vin);
        F = ToMatrix<int64_t>(Fin);
        int64_t const nV = X.cols(), nF = F.cols();
        XVA.assign(static_cast<size_t>(nV), 0.);
        FA.assign(static_cast<size_t>(nF), 0.);
        for (int64_t f = 0; f < nF; ++f)
        {
            double ab[3], ac[3];
            for (int d = 0; d < 3; ++d)
            {
                ab[d] = X(d, F(1, f)) - X(d, F(0, f));
                ac[d] = X(d, F(2, f)) - X(d, F(0, f));
            }
            double const cx = ab[1] * ac[2] - ab[2] * ac[1], cy = ab[2] * ac[0] - ab[0] * ac[2],
                         cz = ab[0] * ac[1] - ab[1] * ac[0];
            double const dbl = std::sqrt(cx * cx + cy * cy + cz * cz);
            for (int k = 0; k < 3; ++k)
                XVA[static_cast<size_t>(F(k, f))] += dbl / 6.;
            FA[static_cast<size_t>(f)] = dbl / 2.;
        }
        return *this;
    }
    /** sim/vbd/Data.cpp:56-60 */
    Data& WithBodies(std::vector<int64_t> Bin)
    {
        B = std::move(Bin);
        return *this;
    }
    /** sim/vbd/Data.cpp:121-127 */
    Data& WithContactParameters(double muCIn, double muFIn, double epsvIn)
    {
        muC = muCIn, muF = muFIn, epsv = epsvIn;
        return *this;
    }
    /** sim/vbd/Data.cpp:129-133 */
    Data& WithActiveSetUpdateFrequency(int32_t frequency)
    {
        activeSetUpdateFrequency = frequency;
        return *this;
    }
    template <class TV>
    Data& WithVelocity(TV const& vin)
    {
        v = ToMatrix<double>(vin);
        return *this;
    }
    template <class TA>
    Data& WithAcceleration(TA const& ain)
    {
        aext = ToMatrix<double>(ain);
        return *this;
    }
    Data& WithMaterial(std::vector<double> rhoeIn, std::vector<double> mueIn, std::vector<double> lambdaeIn)
    {
        rhoe = std::move(rhoeIn);
        mue = std::move(mueIn);
        lambdae = std::move(lambdaeIn);
        return *this;
    }
    Data& WithDirichletConstrainedVertices(std::vector<int64_t> dbcIn)
    {
        dbc = std::move(dbcIn);
        return *this;
    }
    Data& WithVertexColoringStrategy(int orderingIn, int selectionIn)
    {
        ordering = orderingIn;
        selection = selectionIn;
        return *this;
    }
    Data& WithInitializationStrategy(EInitializationStrategy s)
    {
        strategy = s;
        return *this;
    }
    Data& WithRayleighDamping(double kDIn)
    {
        kD = kDIn;
        return *this;
    }
    /** sim/vbd/Data.cpp:141-146 */
    Data& WithChebyshevAcceleration(double rhoIn)
    {
        rho = rhoIn;
        acceleration = PBAT_VBD_ACCELERATION_CHEBYSHEV;
        return *this;
    }
    /** sim/vbd/Data.cpp:148-153 */
    Data& WithAndersonAcceleration(int32_t windowIn)
    {
        window = windowIn;
        acceleration = PBAT_VBD_ACCELERATION_ANDERSON;
        return *this;
    }
    /** sim/vbd/Data.cpp:170-177 */
    Data& WithTrustRegionAcceleration(double etaIn, double tauIn, bool curvedIn = true)
    {
        eta = etaIn;
        tau = tauIn;
        curved = curvedIn;
        acceleration = PBAT_VBD_ACCELERATION_TRUST_REGION;
        return *this;
    }
    Data& WithHessianDeterminantZeroUnder(double zero)
    {
        detHZero = zero;
        return *this;
    }

    /* sim/vbd/Data.cpp:179-244 */
    Data& Construct()
    {
        int64_t const nV = X.cols(), nT = E.cols();
        if (nV == 0 || nT == 0)
            throw std::invalid_argument("WithVolumeMesh must be called with a non-empty mesh");
        x = X;
        if (v.cols() == 0)
            v = Matrix<double>(3, nV);
        if (aext.cols() == 0)
        {
            aext = Matrix<double>(3, nV);
            for (int64_t i = 0; i < nV; ++i)
                aext(2, i) = -9.81;
        }
        if (v.cols() != nV || aext.cols() != nV)
            throw std::invalid_argument("v and aext must be 3 x |#nodes|");
        bool const material = !mue.empty();
        m.assign(static_cast<size_t>(nV), 0.);
        GP = Matrix<double>(4, 3 * nT);
        wg.assign(static_cast<size_t>(nT), 0.);
        lame = Matrix<double>(2, nT);
        GVGp.assign(static_cast<size_t>(nV + 1), 0);
        GVGe.assign(static_cast<size_t>(4 * nT), 0);
        GVGilocal.assign(static_cast<size_t>(4 * nT), 0);
        colors.assign(static_cast<size_t>(nV), 0);
        Pptr.assign(PBAT_VBD_MAX_COLORS + 1, 0);
        Padj.assign(static_cast<size_t>(nV), 0);
        int64_t nP = 0, nAdj = 0;
        Check(pbat_vbd_data_construct(nV, nT, X.data(), E.data(), rhoe.empty() ? nullptr : rhoe.data(),
                                      material ? mue.data() : nullptr, material ? lambdae.data() : nullptr,
                                      dbc.empty() ? nullptr : dbc.data(), static_cast<int64_t>(dbc.size()),
                                      ordering, selection, m.data(), GP.data(), wg.data(), lame.data(),
                                      GVGp.data(), GVGe.data(), GVGilocal.data(), colors.data(), Pptr.data(),
                                      Padj.data(), &nP, &nAdj));
        Pptr.resize(static_cast<size_t>(nP + 1));
        Padj.resize(static_cast<size_t>(nAdj));
        for (int64_t i : dbc)
            for (int d = 0; d < 3; ++d)
                v(d, i) = 0., aext(d, i) = 0.;
        return *this;
    }
};

} // namespace sim::vbd

namespace gpu::vbd {

using GpuMatrixX = Matrix<float>;

/* gpu/vbd/Integrator.h:33-150 */
class Integrator
{
  public:
    using Data = sim::vbd::Data;
    using EInitializationStrategy = sim::vbd::EInitializationStrategy;

    explicit Integrator(Data const& data)
    {
        pbat_vbd_data d{};
        d.nV = data.x.cols();
        d.nT = data.E.cols();
        d.x = data.x.data();
        d.v = data.v.data();
        d.aext = data.aext.data();
        d.m = data.m.data();
        d.E = data.E.data();
        d.GP = data.GP.data();
        d.wg = data.wg.data();
        d.lame = data.lame.data();
        d.GVGp = data.GVGp.data();
        d.GVGe = data.GVGe.data();
        d.GVGilocal = data.GVGilocal.data();
        d.nPartitions = static_cast<int64_t>(data.Pptr.size()) - 1;
        d.Pptr = data.Pptr.data();
        d.Padj = data.Padj.data();
        d.strategy = static_cast<int32_t>(data.strategy);
        d.kD = data.kD;
        d.detHZero = data.detHZero;
        d.acceleration = data.acceleration;
        d.rho = data.rho;
        d.window = data.window;
        d.eta = data.eta;
        d.tau = data.tau;
        d.curved = data.curved ? 1 : 0;
        if (!data.V.empty() && data.F.cols() > 0)
        {
            if (data.B.size() != static_cast<size_t>(d.nV) || data.XVA.size() != static_cast<size_t>(d.nV))
                throw std::invalid_argument("B and XVA must have |#nodes| entries");
            d.nCollisionVertices = static_cast<int64_t>(data.V.size());
            d.V = data.V.data();
            d.nCollisionTriangles = data.F.cols();
            d.F = data.F.data();
            d.B = data.B.data();
            d.XVA = data.XVA.data();
            d.FA = data.FA.data();
        }
        d.muC = data.muC;
        d.muF = data.muF;
        d.epsv = data.epsv;
        d.activeSetUpdateFrequency = data.activeSetUpdateFrequency;
        Check(pbat_vbd_integrator_create(&d, &mImpl));
        mVertices = d.nV;
    }
    Integrator(Integrator const&) = delete;
    Integrator& operator=(Integrator const&) = delete;
    Integrator(Integrator&& o) noexcept : mImpl(o.mImpl), mVertices(o.mVertices) { o.mImpl = nullptr; }
    Integrator& operator=(Integrator&& o) noexcept
    {
        if (this != &o)
        {
            if (mImpl)
                pbat_vbd_integrator_destroy(mImpl);
            mImpl = o.mImpl;
            mVertices = o.mVertices;
            o.mImpl = nullptr;
        }
        return *this;
    }
    ~Integrator()
    {
        if (mImpl)
            pbat_vbd_integrator_destroy(mImpl);
    }

    void Step(float dt, int32_t iterations, int32_t substeps = 1)
    {
        Check(pbat_vbd_integrator_step(mImpl, dt, iterations, substeps));
    }
    void TracedStep(float dt, int32_t iterations, int32_t substeps, int32_t t, std::string const& dir = ".")
    {
        Check(pbat_vbd_integrator_traced_step(mImpl, dt, iterations, substeps, t, dir.c_str()));
    }
    template <class TMatrix>
    void SetPositions(TMatrix const& X)
    {
        auto const M = ToVertexMatrix(X);
        Check(pbat_vbd_integrator_set_positions(mImpl, M.data(), M.cols()));
    }
    template <class TMatrix>
    void SetVelocities(TMatrix const& v)
    {
        auto const M = ToVertexMatrix(v);
        Check(pbat_vbd_integrator_set_velocities(mImpl, M.data(), M.cols()));
    }
    template <class TMatrix>
    void SetExternalAcceleration(TMatrix const& a)
    {
        auto const M = ToVertexMatrix(a);
        Check(pbat_vbd_integrator_set_external_acceleration(mImpl, M.data(), M.cols()));
    }
    void SetNumericalZeroForHessianDeterminant(float zero) { Check(pbat_vbd_integrator_set_detH_zero(mImpl, zero)); }
    void SetRayleighDampingCoefficient(float kD) { Check(pbat_vbd_integrator_set_rayleigh_damping(mImpl, kD)); }
    void SetInitializationStrategy(EInitializationStrategy s)
    {
        Check(pbat_vbd_integrator_set_initialization_strategy(mImpl, static_cast<int32_t>(s)));
    }
    void SetBlockSize(int32_t blockSize) { Check(pbat_vbd_integrator_set_block_size(mImpl, blockSize)); }
    void SetSceneBoundingBox(float const min[3], float const max[3])
    {
        Check(pbat_vbd_integrator_set_scene_bounding_box(mImpl, min, max));
    }
    GpuMatrixX GetPositions() const
    {
        GpuMatrixX M(3, mVertices);
        Check(pbat_vbd_integrator_get_positions(mImpl, M.data(), mVertices));
        return M;
    }
    GpuMatrixX GetVelocities() const
    {
        GpuMatrixX M(3, mVertices);
        Check(pbat_vbd_integrator_get_velocities(mImpl, M.data(), mVertices));
        return M;
    }
    pbat_vbd_integrator* Handle() const { return mImpl; }

  private:
    template <class TMatrix>
    static GpuMatrixX ToVertexMatrix(TMatrix const& A)
    {
        if (A.rows() != 3)
            throw std::invalid_argument("Expected input dimensions 3 x |#vertices|");
        return ToMatrix<float>(A);
    }
    pbat_vbd_integrator* mImpl = nullptr;
    int64_t mVertices = 0;
};

} // namespace gpu::vbd
} // namespace pbat_amd

#endif // PBAT_VBD_HPP
