// param.h -- run parameters of the driver.  Defaults are the reference's
// (param.h:5-28 and the literals of Run::initialize, Run.cpp:14-52): gamma,
// reference Mach number and viscosity, end time, the fixed time step, the
// 2000 x 2000 mesh on a 2 x 2 square, the circle r = 0.1 at (0.5, 1.0) with
// 400 Lagrangian points and 3 ghost layers, 30000 stages, a snapshot every 50.
#pragma once

namespace cfsdrv {

constexpr double Gamma = 1.4;
constexpr double Ma_ref = 1.0;
constexpr double mu_ref = 1.0e-3;
constexpr double T = 1.0;
constexpr double dt = 5.0e-5;

constexpr int N = 2000;
constexpr double LengthX = 2.0, LengthY = 2.0;

constexpr double CircleRadius = 0.1, CircleX = 0.5, CircleY = 1.0;
constexpr int LagrangianPoints = 400, GhostLayers = 3;

constexpr int Stages = 30000;
constexpr int OutputEvery = 50;

}  // namespace cfsdrv
