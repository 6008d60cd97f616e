// square_driver.cpp -- a tspeed-style driver (same shape as the reference's
// Examples/src/Lamb.cpp) running the 1 km square workload through the tspeed::LeapFrog facade:
// 1000 m x 1000 m, 40x40 split squares, p = 4, vertical Ricker at (500, 950),
// 32 receivers at y = 990.  Writes <out>_rcv_<i>.out.
//
//   build:  g++ -O2 -std=c++17 -I../tspeed_amd/csrc/host square_driver.cpp \
//               -L../tspeed_amd -ltspeed_amd -Wl,-rpath,$ORIGIN/../tspeed_amd -o square_driver
//   run:    ./square_driver [steps] [out_prefix] [precision 0|1] [basis 0|1] [digits]
//           basis 1 = FESpace<N, Dunavant<2N>, BoundaryAdapted<N>> (as Lamb.cpp:27);
//           digits 0 (default) writes the reference's 6-digit files, 17 every bit
#include "tspeed.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

using namespace tspeed;

int main(int argc, char **argv)
{
    const int nsteps = argc > 1 ? std::atoi(argv[1]) : 2000;
    const std::string out = argc > 2 ? argv[2] : "square";
    const int precision = argc > 3 ? std::atoi(argv[3]) : 0;
    const int basis = argc > 4 ? std::atoi(argv[4]) : 0;
    const int digits = argc > 5 ? std::atoi(argv[5]) : 0;
    const int N = 4;

    Mesh_ptr Th = Mesh::split_square(40, 40, 0.0, 0.0, 1000.0, 1000.0, 0.2, 1);
    FESpace_ptr Xh = basis ? FESpace_ptr(new FESpace<N, Dunavant<2 * N>, BoundaryAdapted<N>>(Th))
                           : FESpace_ptr(new FESpace<N>(Th));

    Parameters p(Th);
    const double rho = 2000, vp = 3000, vs = 1732;
    p.setp("rho", 0, rho);
    p.setp("mu", 0, rho * vs * vs);
    p.setp("lambda", 0, rho * vp * vp - 2 * rho * vs * vs);

    std::vector<Geo::Point> pts;
    for (int i = 0; i < 32; ++i) pts.emplace_back((i + 0.5) * 1000.0 / 32, 990.0);
    Receivers r(Xh, pts);

    const double f0 = 10, t0 = 1.2 / f0;
    Force_ptr f(new PointWiseForce([=](const double &t) { return std::array<double, 2>{{0, -ricker(f0, t0, t)}}; },
                                   Geo::Point(500, 950), Xh));

    const double dt = 0.2 * Th->min_edge() / (vp * N * N);
    LeapFrog TA(Xh, p, r, /*penalty=*/10.0, precision);
    TA.set_dt(dt);
    TA.set_tmax((nsteps + 2) * dt);
    TA.add_force(f);
    TA.set_initial_u(Xh, [](double, double) { return std::array<double, 2>{{0, 0}}; });
    TA.set_initial_v(Xh, [](double, double) { return std::array<double, 2>{{0, 0}}; });

    TA.first_step();
    TA.eval_receivers();
    // reference-style loop: one step() per iteration, sampling every step
    int step = 0;
    while (TA.is_running() && step < nsteps) {
        TA.step();
        ++step;
        TA.eval_receivers();
    }
    TA.write_receivers(out, digits);
    const auto &u = TA.get_uh();
    double nrm = 0;
    for (double v : u) nrm += v * v;
    std::printf("square_driver: %d steps, dt=%.6e, |u|=%.17e\n", step, dt, std::sqrt(nrm));
    return 0;
}
