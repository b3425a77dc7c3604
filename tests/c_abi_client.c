/* A C99 client of the drop-in boundary: includes every public header, links against
 * libplens_gpu.so and calls the host-side entry points (no GPU needed). Prints "ok" and returns 0
 * when every check passes. Built and run by tests/test_c_abi_client.py. */
#include <math.h>
#include <stdio.h>

#include "plens_gpu.h"
#include "plens_host.h"
#include "plens_mlp.h"

#define CHECK(cond)                                                   \
    do {                                                              \
        if (!(cond)) {                                                \
            printf("check failed at line %d: %s\n", __LINE__, #cond); \
            return 1;                                                 \
        }                                                             \
    } while (0)

int main(void) {
    /* face-local dof -> cell node, N = 2 (dgsem/face_dof_info.h:84-89) */
    static const int expect[6][9] = {{0, 3, 6, 9, 12, 15, 18, 21, 24}, {2, 5, 8, 11, 14, 17, 20, 23, 26},
                                     {0, 1, 2, 9, 10, 11, 18, 19, 20}, {6, 7, 8, 15, 16, 17, 24, 25, 26},
                                     {0, 1, 2, 3, 4, 5, 6, 7, 8},      {18, 19, 20, 21, 22, 23, 24, 25, 26}};
    for (int f = 0; f < 6; ++f)
        for (int q = 0; q < 9; ++q) CHECK(plens_face_node(2, f, q) == expect[f][q]);

    /* GLL basis of degree 4 on [0,1]: symmetric nodes, weights summing to 1, D annihilates constants */
    double x[5], w[5], D[25], C1[25];
    CHECK(plens_basis_1d(4, x, w, D, C1) == 0);
    double ws = 0;
    for (int i = 0; i < 5; ++i) {
        ws += w[i];
        CHECK(fabs(x[i] + x[4 - i] - 1.0) < 1e-15);
        double rs = 0;
        for (int j = 0; j < 5; ++j) rs += D[i * 5 + j];
        CHECK(fabs(rs) < 1e-12);
    }
    CHECK(fabs(ws - 1.0) < 1e-15);

    /* varying-inflow expressions: three components, muParser precedence (^ before unary minus) */
    double v[3];
    CHECK(plens_expr_eval("x + 2*t; -2^2; if(y > 1, 3, 4)", 3, 1.0, 2.0, 0.0, 0.5, v) == 0);
    CHECK(v[0] == 2.0 && v[1] == -4.0 && v[2] == 3.0);
    CHECK(plens_expr_eval("x +", 1, 0, 0, 0, 0, v) < 0);

    /* conservative state from primitive variables */
    const double vel[3] = {1.0, 2.0, 0.0};
    double c[5];
    plens_prim_to_cons(1.4, 2.0, vel, 1.0, c);
    CHECK(c[0] == 2.0 && c[1] == 2.0 && c[2] == 4.0 && fabs(c[4] - (2.5 + 5.0)) < 1e-15);

    /* a box mesh and its descriptor, the input of plens_gpu_create */
    plens_box_params prm = {0};
    prm.degree = 3;
    prm.nx = 4; prm.ny = 3; prm.nz = 2;
    prm.hi[0] = 1.0; prm.hi[1] = 1.0; prm.hi[2] = 1.0;
    prm.y_stretch_ratio = 1.0;
    prm.periodic[0] = prm.periodic[1] = prm.periodic[2] = 1;
    plens_mesh* m = NULL;
    CHECK(plens_mesh_build_box(&prm, &m) == 0 && m);
    CHECK(plens_mesh_n_cells(m) == 24 && plens_mesh_degree(m) == 3);
    plens_mesh_desc desc;
    CHECK(plens_mesh_fill_desc(m, &desc) == 0);
    CHECK(desc.n_cells == 24 && desc.degree == 3 && desc.face_nbr && desc.face_bid);
    for (int cf = 0; cf < 24 * 6; ++cf) CHECK(desc.face_nbr[cf] >= 0 && desc.face_nbr[cf] < 24);
    plens_mesh_destroy(m);

    /* the network's parameter count: W1[64][125], b1, W2[64][64], b2, w3, b3 */
    CHECK(plens_mlp_n_params() == 64 * 125 + 64 + 64 * 64 + 64 + 64 + 1);

    /* argument checks of the device entry points that precede any device work */
    CHECK(plens_gpu_create(NULL, NULL, NULL, 0, NULL) != 0);
    printf("ok\n");
    return 0;
}
