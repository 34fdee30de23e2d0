// Opaque C-ABI handle types shared by the entry-point translation units.
#pragma once
#include "mesh.hpp"
#include "solver.hpp"

struct tmi_mesh_s {
    tmi::Mesh m;
};

struct tmi_solver_s {
    tmi::Solver s;
};
