// constitutive.h -- host side of the constitutive plugin of the explicit path.
//
// Mirrors the reference's class constitutiveModels (constitutiveModels.h:16-96)
// for the laws with a device descriptor: the two neo-Hookean laws, St-Venant-Kirchhoff
// and linear elastic.  The per-Gauss-point update
// (updateConstitutive, defoEnergy) runs in the device element kernels; the
// host object keeps the law's name, parameters and sound speed, and hands the
// kernels a device descriptor (law id + E, nu, rho).  A law without a
// descriptor cannot be built: the reader rejects it.
#ifndef MPS_CONSTITUTIVE_H
#define MPS_CONSTITUTIVE_H

#include <memory>
#include <string>

namespace mps {

struct DeviceMaterial {
    int law; // MPS_LAW_*
    double young, nu, rho;
};

class constitutiveModels {
  public:
    constitutiveModels(int ndim, std::string name, std::string nameCons, double rho)
        : _ndim(ndim), _name(std::move(name)), _nameCons(std::move(nameCons)), _rho(rho) {}
    virtual ~constitutiveModels() = default;

    const std::string &getName() const { return _name; }         // constitutiveModels.h:30-36
    const std::string &getNameCons() const { return _nameCons; } // keyword after "*HyperElastic, "
    double getDensity() const { return _rho; }
    int getNbrDofConsMod() const { return _ndim; } // mechanical DOFs only
    virtual double soundSpeed() const = 0;
    virtual DeviceMaterial deviceDescriptor() const = 0;

  protected:
    int _ndim;
    std::string _name, _nameCons;
    double _rho;
};

// "*HyperElastic, Neo-Hookean" (hyperElasticNeoHookean.cpp:23-164):
// sigma = mu J^{-5/3} dev(B) + K (J-1) I, K = E / (3(1-2nu)).
class classHyperElasticNeoHookean : public constitutiveModels {
  public:
    classHyperElasticNeoHookean(int ndim, std::string name, std::string nameCons, double rho, double young, double nu);
    double soundSpeed() const override; // sqrt((lambda + 2 mu)/rho), hyperElasticNeoHookean.cpp:51-54
    DeviceMaterial deviceDescriptor() const override;
    double getLambda() const { return _lambda; }
    double getMu() const { return _mu; }
    double getK() const { return _K1; }

  private:
    double _Young, _nu, _lambda, _mu, _K1;
};

// "*HyperElastic, 3D-Compressible-Neo-Hookean" (hyperElasticNeoHookean.cpp:170-304):
// S = mu I + (lambda lnJ - mu) C^{-1}, P = F S, W = lambda/2 ln^2 J - mu lnJ + mu/2 tr C.
class classHyperElastic3DCompressibleNeoHookean : public constitutiveModels {
  public:
    classHyperElastic3DCompressibleNeoHookean(int ndim, std::string name, std::string nameCons, double rho,
                                              double young, double nu);
    double soundSpeed() const override; // hyperElasticNeoHookean.cpp:202-205
    DeviceMaterial deviceDescriptor() const override;
    double getLambda() const { return _lambda; }
    double getMu() const { return _mu; }

  private:
    double _Young, _nu, _lambda, _mu;
};

// "*HyperElastic, St-Venant-Kirchhoff" (hyperElasticStVenantKirchhoff.cpp): S = lambda tr(E) I + 2 mu E,
// P = F S, E the Green-Lagrange strain; sound speed sqrt(E/rho).
class classHyperElasticStVenantKirchhoff : public constitutiveModels {
  public:
    classHyperElasticStVenantKirchhoff(int ndim, std::string name, std::string nameCons, double rho, double young,
                                       double nu)
        : constitutiveModels(ndim, std::move(name), std::move(nameCons), rho), _Young(young), _nu(nu) {}
    double soundSpeed() const override;
    DeviceMaterial deviceDescriptor() const override;

  private:
    double _Young, _nu;
};

// "*Linear-Elastic" (linearElastic.cpp): P = lambda tr(e) I + 2 mu e, e = sym(F) - I; sound speed sqrt(E/rho).
class linearElastic : public constitutiveModels {
  public:
    linearElastic(int ndim, std::string name, std::string nameCons, double rho, double young, double nu)
        : constitutiveModels(ndim, std::move(name), std::move(nameCons), rho), _Young(young), _nu(nu) {}
    double soundSpeed() const override;
    DeviceMaterial deviceDescriptor() const override;

  private:
    double _Young, _nu;
};

// Factory keyed by the law keyword, as materialManagement does
// (inputFile.cpp:1008-1014).  Returns null for a law without a descriptor.
std::unique_ptr<constitutiveModels> makeConstitutiveModel(const std::string &name, const std::string &nameCons,
                                                          double rho, double young, double nu);

} // namespace mps

#endif
