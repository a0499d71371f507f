// constitutive.cpp -- host side of the neo-Hookean laws (constitutive.h).
#include "constitutive.h"

#include <cmath>

#include "mps_gpu.h"

namespace mps {

// Lame constants as the reference constructors form them (hyperElasticNeoHookean.cpp:23-31, 175-182).
classHyperElasticNeoHookean::classHyperElasticNeoHookean(int ndim, std::string name, std::string nameCons, double rho,
                                                         double young, double nu)
    : constitutiveModels(ndim, std::move(name), std::move(nameCons), rho), _Young(young), _nu(nu),
      _lambda((young * nu) / ((1. + nu) * (1. - 2. * nu))), _mu(0.5 * young / (1. + nu)),
      _K1(young / (3 * (1 - 2 * nu))) {}

double classHyperElasticNeoHookean::soundSpeed() const { return std::sqrt((_lambda + 2 * _mu) / _rho); }

DeviceMaterial classHyperElasticNeoHookean::deviceDescriptor() const {
    return DeviceMaterial{MPS_LAW_NH_BELYTSCHKO, _Young, _nu, _rho};
}

classHyperElastic3DCompressibleNeoHookean::classHyperElastic3DCompressibleNeoHookean(int ndim, std::string name,
                                                                                     std::string nameCons, double rho,
                                                                                     double young, double nu)
    : constitutiveModels(ndim, std::move(name), std::move(nameCons), rho), _Young(young), _nu(nu),
      _lambda((young * nu) / ((1. + nu) * (1. - 2. * nu))), _mu(0.5 * young / (1. + nu)) {}

double classHyperElastic3DCompressibleNeoHookean::soundSpeed() const { return std::sqrt((_lambda + 2 * _mu) / _rho); }

DeviceMaterial classHyperElastic3DCompressibleNeoHookean::deviceDescriptor() const {
    return DeviceMaterial{MPS_LAW_NH_COMPRESSIBLE, _Young, _nu, _rho};
}

double classHyperElasticStVenantKirchhoff::soundSpeed() const { return std::sqrt(_Young / _rho); }

DeviceMaterial classHyperElasticStVenantKirchhoff::deviceDescriptor() const {
    return DeviceMaterial{MPS_LAW_SVK, _Young, _nu, _rho};
}

double linearElastic::soundSpeed() const { return std::sqrt(_Young / _rho); }  // linearElastic.cpp:49-54

DeviceMaterial linearElastic::deviceDescriptor() const { return DeviceMaterial{MPS_LAW_LINEAR_ELASTIC, _Young, _nu, _rho}; }

std::unique_ptr<constitutiveModels> makeConstitutiveModel(const std::string &name, const std::string &nameCons,
                                                          double rho, double young, double nu) {
    if (nameCons == "Neo-Hookean")
        return std::unique_ptr<constitutiveModels>(new classHyperElasticNeoHookean(3, name, nameCons, rho, young, nu));
    if (nameCons == "3D-Compressible-Neo-Hookean")
        return std::unique_ptr<constitutiveModels>(
            new classHyperElastic3DCompressibleNeoHookean(3, name, nameCons, rho, young, nu));
    if (nameCons == "St-Venant-Kirchhoff")
        return std::unique_ptr<constitutiveModels>(
            new classHyperElasticStVenantKirchhoff(3, name, nameCons, rho, young, nu));
    if (nameCons == "Linear-Elastic")
        return std::unique_ptr<constitutiveModels>(new linearElastic(3, name, nameCons, rho, young, nu));
    return nullptr;
}

} // namespace mps
