// Lattice images of the reference's translational sums, for the adapters
// (header-only; depends on lattice_point.h alone).
#ifndef HIP_LATTICE_H
#define HIP_LATTICE_H

#include <vector>

#include "lattice_point.h"

namespace hip_adapter {

// The signed lattice images of a Grid in the order the reference's
// translational sums visit them (direct_force_method_translational.cpp:28-309,
// barnes_hut_force_calculation_translational.cpp:27-178): each point of
// grid.lattice expanded over its sign and axis permutations, 3 ints per image.
inline int oddEven(int n) { return (n & 1) ? -1 : 1; }

inline std::vector<int> expandLattice(const Grid& grid) {
    std::vector<int> out;
    auto push = [&out](int x, int y, int z) {
        out.push_back(x);
        out.push_back(y);
        out.push_back(z);
    };
    const int ax = grid.symmetryAxes.x, ay = grid.symmetryAxes.y, az = grid.symmetryAxes.z;
    for (int ii = 0; ii < grid.numLattice; ii++) {
        const int px = grid.lattice[ii].point.x, py = grid.lattice[ii].point.y, pz = grid.lattice[ii].point.z;
        const int nz = (px != 0) + (py != 0) + (pz != 0);
        const int rx = 1 + (px != 0), ry = 1 + (py != 0), rz = 1 + (pz != 0);
        if (grid.symDimSum == 3 && nz == 3) {
            for (int i = 0; i < 2; i++)
                for (int j = 0; j < 2; j++)
                    for (int k = 0; k < 2; k++) push(oddEven(i) * px, oddEven(j) * py, oddEven(k) * pz);
        } else if (grid.symDimSum == 3 && nz == 2) {
            for (int jj = 0; jj < 3; jj++)
                for (int i = 0; i < rx; i++)
                    for (int j = 0; j < ry; j++)
                        for (int k = 0; k < rz; k++) {
                            const int oi = oddEven(i), oj = oddEven(j), ok = oddEven(k);
                            push((jj != 0) * oi * px + (px == 0) * ((jj == 1) * oj * py + (jj == 2) * ok * pz),
                                 (jj != 1) * oj * py + (py == 0) * ((jj == 0) * oi * px + (jj == 2) * ok * pz),
                                 (jj != 2) * ok * pz + (pz == 0) * ((jj == 0) * oi * px + (jj == 1) * oj * py));
                        }
        } else if (grid.symDimSum == 3 && nz == 1) {
            for (int jj = 0; jj < 3; jj++)
                for (int i = 0; i < rx; i++)
                    for (int j = 0; j < ry; j++)
                        for (int k = 0; k < rz; k++) {
                            const int oi = oddEven(i), oj = oddEven(j), ok = oddEven(k);
                            push((jj == 0) * oi * px + (jj == 1) * oj * py + (jj == 2) * ok * pz,
                                 (jj == 0) * oj * py + (jj == 1) * ok * pz + (jj == 2) * oi * px,
                                 (jj == 0) * ok * pz + (jj == 1) * oi * px + (jj == 2) * oj * py);
                        }
        } else if (grid.symDimSum == 2 && nz == 2) {
            for (int i = 0; i < rx; i++)
                for (int j = 0; j < ry; j++)
                    for (int k = 0; k < rz; k++) push(oddEven(i) * px, oddEven(j) * py, oddEven(k) * pz);
        } else if (grid.symDimSum == 2 && nz == 1) {
            for (int jj = 0; jj < 2; jj++)
                for (int i = 0; i < rx; i++)
                    for (int j = 0; j < ry; j++)
                        for (int k = 0; k < rz; k++) {
                            const int oi = oddEven(i), oj = oddEven(j), ok = oddEven(k);
                            push(ax * ((jj == 0) * oi * px + (jj == 1) * (oj * py + ok * pz)),
                                 ay * ((jj == 0) * oj * py + (jj == 1) * (oi * px + ok * pz)),
                                 az * ((jj == 0) * ok * pz + (jj == 1) * (oi * px + oj * py)));
                        }
        } else if (grid.symDimSum == 1 && nz == 1) {
            for (int i = 0; i < rx; i++)
                for (int j = 0; j < ry; j++)
                    for (int k = 0; k < rz; k++) push(oddEven(i) * px, oddEven(j) * py, oddEven(k) * pz);
        } else {
            push(0, 0, 0);
        }
    }
    return out;
}

}  // namespace hip_adapter

#endif  // HIP_LATTICE_H
