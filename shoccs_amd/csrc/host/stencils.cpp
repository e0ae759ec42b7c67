// stencils.cpp -- interior stencils and numerical boundary schemes.
//
// Closed-form closure coefficients of the reference's symbolic generator,
// evaluated in the reference's operation order (the product is built with
// -ffp-contract=off) so that every coefficient is the reference's double:
//
//   E4_1  (P2 R5 T7, cut-cell)  src/stencils/E4_1.cpp:43-342
//   E4u_1 (P2 R3 T5)            src/stencils/E4u_1.cpp:26-126
//   E4_2  (P2 R3 T5 X3)         src/stencils/E4_2.cpp:16-42,185-296
//   E2_2  (P1 R2 T4 X2)         src/stencils/E2_2.cpp:12-180
//   E2-poly (P1 R3 T4)          src/stencils/polyE2_1.cpp:10-257
//
// Right-wall closures are the left ones reversed (negated for the first
// derivative).
#include "stencils.hpp"

#include <algorithm>
#include <cmath>

namespace ccs::stencils {

namespace {

using std::pow;

void reverse(double* c, int n) { std::reverse(c, c + n); }

double clampd(double v, double lo, double hi) { return std::clamp(v, lo, hi); }

/* ---------------------------------------------------------------- E4_1 */
void e41_floating(const double* alpha, double h, double psi, double* c, int right)
{
    const double psi_eps = 1e-4;
    psi = clampd(psi, psi_eps, 1.0 - psi_eps);
    const double a0 = alpha[0], a1 = alpha[1];

    double t5 = a1 * psi;
    double t6 = psi * psi;
    double t7 = psi * psi * psi;
    double t8 = pow(psi, 4);
    double t9 = pow(psi, 5);
    double t10 = pow(psi, 6);
    double t11 = pow(psi, 7);
    double t12 = a1 * t6;
    double t13 = a1 * t7;
    double t14 = a1 * t8;
    double t15 = a1 * t9;
    double t16 = a1 * t10;
    double t17 = a1 * t11;
    double t18 = a1 * pow(psi, 8);
    double t19 = a1 * a1;
    double t20 = psi * t19;
    double t21 = a0 * t12;
    double t22 = a0 * t13;
    double t23 = a0 * t14;
    double t24 = a0 * t15;
    double t25 = a0 * t16;
    double t26 = t19 * t6;
    double t27 = 99360 * t19;
    double t28 = t19 * t9;
    double t29 = a0 * t26;
    double t30 = a0 * t19;
    double t31 = 17280 * t30;
    double t32 = 2484 * a0 * t17 + 216 * a0 * t18 + 3456 * a0 * t28 + 10368 * a1 + 13872 * psi -
                 414 * t10 - 36 * t11 + 314677 * t12 + 1282709 * t13 - 182437 * t14 -
                 496397 * t15 - 169230 * t16 - 21558 * t17 - 1716 * t18 + 276480 * t20 +
                 20394 * t21 - 94806 * t22 + 14166 * t23 + 41670 * t24 + 15876 * t25 +
                 218592 * t26 - t27 * t7 - t27 * t8 - 19872 * t28 - 38016 * t29 + t31 * t7 +
                 t31 * t8 - 184800 * t5 + 33753 * t6 + 9060 * t7 - 6339 * t8 - 2988 * t9 - 7092;
    double t33 = 1.0 / (a1);
    double t34 = 1.0 / (psi - 1);
    double t35 = 1.0 / (1728 * a1 + 1721 * psi + 12 * t10 + 864 * t12 + 144 * t13 + 1584 * t5 +
                        8070 * t6 + 5617 * t7 + 1464 * t8 + 162 * t9 - 1182);
    double t36 = t33 * t34 * t35;
    double t37 = t32 * t36;
    double t38 = (1.0 / 6) * t37;
    double t39 = 1.0 / (psi);
    double t40 = t38 * t39;
    double t41 = -t40;
    double t42 = -2.0 / 3 * t32 * t33 * t34 * t35 * t39;
    double t43 = t37 * t6;
    double t44 = psi * t37;
    double t45 = t37 * t7;
    double t46 = (1.0 / 36) * t45 + (11.0 / 36) * t8;
    double t47 = (7.0 / 12) * t6;
    double t48 = (1.0 / 12) * t45 + (11.0 / 12) * t8;
    double t49 = (5.0 / 36) * t6;
    double t50 = a0 * psi;
    double t51 = 6 * a0;
    double t52 = 168 * a0 * a1 - 832 * a1 * t50 + 342 * a1 - 288 * psi - 10194 * t12 +
                 4350 * t13 + 4671 * t14 + 1032 * t15 + 33 * t16 + t17 * t51 + 3 * t17 -
                 756 * t19 * t50 + 330 * t19 * t7 + 66 * t19 * t8 + 5034 * t20 - 770 * t21 +
                 274 * t22 + 752 * t23 + 336 * t24 + 66 * t25 + 330 * t26 + 540 * t29 +
                 540 * t30 * t7 + 108 * t30 * t8 - 432 * t30 - 5565 * t5 - 18 * t6 + 72 * t7 +
                 18 * t8 - 216;
    double t53 = t36 * t52;
    double t54 = 4 * t53;
    double t55 = psi * t54;
    double t56 = (1.0 / 3) * psi;
    double t57 = t56 - 1.0 / 3;
    double t58 = t53 * t8;
    double t59 = -2.0 / 3 * t58;
    double t60 = 16 * t53;
    double t61 = psi * t53;
    double t62 = (1.0 / 18) * t8;
    double t63 = -13.0 / 9 * psi + (17.0 / 18) * t6 + t62 + (4.0 / 9) * t7;
    double t64 = 24 * t53;
    double t65 = t53 * t6;
    double t66 = t53 * t7;
    double t67 = -2 * t58;
    double t68 = (1.0 / 6) * t8;
    double t69 = -19.0 / 6 * psi + (11.0 / 6) * t6 + t68 + (7.0 / 6) * t7;
    double t70 = -7.0 / 3 * psi + (7.0 / 6) * t6 + t68 + t7;
    double t71 = -11.0 / 18 * psi + (5.0 / 18) * t6 + t62 + (5.0 / 18) * t7;
    double t72 = -a0;
    double t73 = -4 * a0;
    double t74 = a0 * t7;
    double t75 = a0 * t6;
    double t76 = a0 * t68 + (1.0 / 36) * t8;
    double t77 = (1.0 / 2) * a0 * t8 + (1.0 / 12) * t8;
    double t78 = 6 * t50;
    double t79 = psi * t64;
    double t80 = -1.0 / 6 * t32 * t33 * t34 * t35;
    double t81 = 2 * psi;
    double t82 = (4.0 / 3) * t37;
    double t83 = 8 * psi;
    double t84 = t51 * t6 + t6 * t64 - 149 * t6;

    c[0] = (11.0 / 6) * psi + t38 + t41 - 11.0 / 6;
    c[1] = psi * (-t40 - 11.0 / 6);
    c[2] = (143.0 / 18) * psi + (13.0 / 18) * t32 * t33 * t34 * t35 - t42 - 2.0 / 9 * t43 -
           17.0 / 36 * t44 - t46 - 187.0 / 36 * t6 - 22.0 / 9 * t7 + 3;
    c[3] = -209.0 / 12 * psi - t37 * t39 + t37 * t47 - 19.0 / 12 * t37 + (11.0 / 12) * t44 +
           t48 + (121.0 / 12) * t6 + (77.0 / 12) * t7 - 3.0 / 2;
    c[4] = (77.0 / 6) * psi + (7.0 / 6) * t32 * t33 * t34 * t35 - t42 - 1.0 / 2 * t43 -
           7.0 / 12 * t44 - t48 - 77.0 / 12 * t6 - 11.0 / 2 * t7 + 1.0 / 3;
    c[5] = -121.0 / 36 * psi + t37 * t49 - 11.0 / 36 * t37 + t41 + (5.0 / 36) * t44 + t46 +
           (55.0 / 36) * t6 + (55.0 / 36) * t7;
    c[6] = 0;
    c[7] = t54 - t55 + t57;
    c[8] = psi * (t54 - 1.0 / 3);
    c[9] = (34.0 / 3) * t33 * t34 * t35 * t52 * t6 + (16.0 / 3) * t33 * t34 * t35 * t52 * t7 -
           t59 - t60 - 52.0 / 3 * t61 - t63 - 1.0 / 2;
    c[10] = 38 * t61 + t64 - 22 * t65 - 14 * t66 + t67 + t69 + 1;
    c[11] = 14 * t33 * t34 * t35 * t52 * t6 + 12 * t33 * t34 * t35 * t52 * t7 - t60 - 28 * t61 -
            t67 - t70 - 1.0 / 6;
    c[12] = t54 + t59 + (22.0 / 3) * t61 - 10.0 / 3 * t65 - 10.0 / 3 * t66 + t71;
    c[13] = 0;
    c[14] = -1.0 / 6 * psi - t50 - t72 + 1.0 / 6;
    c[15] = psi * (a0 + 1.0 / 6);
    c[16] = -13.0 / 18 * psi - 13.0 / 3 * t50 + (17.0 / 36) * t6 + (2.0 / 9) * t7 + t73 +
            (4.0 / 3) * t74 + (17.0 / 6) * t75 + t76 - 1;
    c[17] = (19.0 / 2) * a0 * psi + (19.0 / 12) * psi + t51 - 11.0 / 12 * t6 - 7.0 / 12 * t7 -
            7.0 / 2 * t74 - 11.0 / 2 * t75 - t77 + 1.0 / 2;
    c[18] = -7.0 / 6 * psi + t47 - 7 * t50 + (1.0 / 2) * t7 + t73 + 3 * t74 + (7.0 / 2) * t75 +
            t77 + 1.0 / 3;
    c[19] = (11.0 / 6) * a0 * psi + (11.0 / 36) * psi - t49 - 5.0 / 36 * t7 - t72 -
            5.0 / 6 * t74 - 5.0 / 6 * t75 - t76;
    c[20] = 0;
    c[21] = t57;
    c[22] = -t56;
    c[23] = 3.0 / 2 - t63;
    c[24] = t69 - 3;
    c[25] = 11.0 / 6 - t70;
    c[26] = t71;
    c[27] = 0;
    c[28] = (-149 * psi - t51 - t64 + t78 + t79 - 11) / (11 * psi + 6 * t6 + t7 + 6);
    c[29] = (149.0 / 6) * psi - t50 - t55 - t80;
    c[30] = (8 * a0 - psi * t82 + 3 * psi + 2 * t50 + t53 * t83 + 32 * t53 - t82 + t84 + 14) /
            (t81 + 2);
    c[31] = (-24 * a0 + 2 * psi * t32 * t33 * t34 * t35 + 4 * t32 * t33 * t34 * t35 -
             96 * t53 - t78 - t79 - t83 - t84 - 38) /
            (t81 + 4);
    c[32] = (72 * a0 - 4 * psi * t37 + 15 * psi - 12 * t37 + 18 * t50 + 288 * t53 + 72 * t61 +
             t84 + 78) /
            (6 * psi + 18);
    c[33] = -a0 - t54 - t80;
    c[34] = 0;

    const int n = 35;
    for (int i = 0; i < n; i++) c[i] /= h;
    if (right) {
        for (int i = 0; i < n; i++) c[i] *= -1;
        reverse(c, n);
    }
}

void e41_dirichlet(const double* alpha, double h, double psi, double* c, int right)
{
    const double psi_eps = 1e-4;
    psi = clampd(psi, psi_eps, 1.0 - psi_eps);
    const double a0 = alpha[0], a1 = alpha[1];

    double t5 = a0 * a1;
    double t6 = a1 * psi;
    double t7 = psi * psi;
    double t8 = psi * psi * psi;
    double t9 = pow(psi, 4);
    double t10 = a0 * psi;
    double t11 = a1 * a1;
    double t12 = a0 * t11;
    double t13 = a1 * t7;
    double t14 = a1 * t8;
    double t15 = a1 * t9;
    double t16 = pow(psi, 5);
    double t17 = a1 * t16;
    double t18 = pow(psi, 6);
    double t19 = a1 * t18;
    double t20 = pow(psi, 7);
    double t21 = a1 * t20;
    double t22 = psi * t11;
    double t23 = a0 * t13;
    double t24 = a0 * t14;
    double t25 = t5 * t9;
    double t26 = t16 * t5;
    double t27 = a0 * t19;
    double t28 = 6 * a0;
    double t29 = 330 * t11;
    double t30 = t11 * t9;
    double t31 = 540 * t12;
    double t32 = t12 * t9;
    double t33 = -832 * a1 * t10 + 342 * a1 - 288 * psi - 756 * t10 * t11 - 432 * t12 -
                 10194 * t13 + 4350 * t14 + 4671 * t15 + 1032 * t17 + 33 * t19 + t21 * t28 +
                 3 * t21 + 5034 * t22 - 770 * t23 + 274 * t24 + 752 * t25 + 336 * t26 +
                 66 * t27 + t29 * t7 + t29 * t8 + 66 * t30 + t31 * t7 + t31 * t8 + 108 * t32 +
                 168 * t5 - 5565 * t6 - 18 * t7 + 72 * t8 + 18 * t9 - 216;
    double t34 = 1.0 / (a1);
    double t35 = 1.0 / (psi - 1);
    double t36 = 1.0 / (1728 * a1 + 1721 * psi + 864 * t13 + 144 * t14 + 162 * t16 + 12 * t18 +
                        1584 * t6 + 8070 * t7 + 5617 * t8 + 1464 * t9 - 1182);
    double t37 = t34 * t35 * t36;
    double t38 = t33 * t37;
    double t39 = 4 * t38;
    double t40 = psi * t39;
    double t41 = (1.0 / 3) * psi;
    double t42 = t41 - 1.0 / 3;
    double t43 = t38 * t9;
    double t44 = -2.0 / 3 * t43;
    double t45 = 16 * t38;
    double t46 = psi * t38;
    double t47 = (1.0 / 18) * t9;
    double t48 = -13.0 / 9 * psi + t47 + (17.0 / 18) * t7 + (4.0 / 9) * t8;
    double t49 = 24 * t38;
    double t50 = t38 * t7;
    double t51 = t38 * t8;
    double t52 = -2 * t43;
    double t53 = (1.0 / 6) * t9;
    double t54 = -19.0 / 6 * psi + t53 + (11.0 / 6) * t7 + (7.0 / 6) * t8;
    double t55 = -7.0 / 3 * psi + t53 + (7.0 / 6) * t7 + t8;
    double t56 = -11.0 / 18 * psi + t47 + (5.0 / 18) * t7 + (5.0 / 18) * t8;
    double t57 = -a0;
    double t58 = -4 * a0;
    double t59 = a0 * t8;
    double t60 = a0 * t7;
    double t61 = a0 * t53 + (1.0 / 36) * t9;
    double t62 = (1.0 / 2) * a0 * t9 + (1.0 / 12) * t9;
    double t63 = 6 * t10;
    double t64 = psi * t49;
    double t65 = pow(psi, 8);
    double t66 = 2484 * a0 * t21 - 1716 * a1 * t65 + 10368 * a1 + 13872 * psi -
                 19872 * t11 * t16 + 218592 * t11 * t7 - 99360 * t11 * t8 + 3456 * t12 * t16 -
                 38016 * t12 * t7 + 17280 * t12 * t8 + 314677 * t13 + 1282709 * t14 -
                 182437 * t15 - 2988 * t16 - 496397 * t17 - 414 * t18 - 169230 * t19 -
                 36 * t20 - 21558 * t21 + 276480 * t22 + 20394 * t23 - 94806 * t24 +
                 14166 * t25 + 41670 * t26 + 15876 * t27 - 99360 * t30 + 17280 * t32 +
                 216 * t5 * t65 - 184800 * t6 + 33753 * t7 + 9060 * t8 - 6339 * t9 - 7092;
    double t67 = -1.0 / 6 * t34 * t35 * t36 * t66;
    double t68 = 2 * psi;
    double t69 = t37 * t66;
    double t70 = (4.0 / 3) * t69;
    double t71 = 8 * psi;
    double t72 = t28 * t7 + t49 * t7 - 149 * t7;

    c[0] = t39 - t40 + t42;
    c[1] = psi * (t39 - 1.0 / 3);
    c[2] = (34.0 / 3) * t33 * t34 * t35 * t36 * t7 + (16.0 / 3) * t33 * t34 * t35 * t36 * t8 -
           t44 - t45 - 52.0 / 3 * t46 - t48 - 1.0 / 2;
    c[3] = 38 * t46 + t49 - 22 * t50 - 14 * t51 + t52 + t54 + 1;
    c[4] = 14 * t33 * t34 * t35 * t36 * t7 + 12 * t33 * t34 * t35 * t36 * t8 - t45 - 28 * t46 -
           t52 - t55 - 1.0 / 6;
    c[5] = t39 + t44 + (22.0 / 3) * t46 - 10.0 / 3 * t50 - 10.0 / 3 * t51 + t56;
    c[6] = 0;
    c[7] = -1.0 / 6 * psi - t10 - t57 + 1.0 / 6;
    c[8] = psi * (a0 + 1.0 / 6);
    c[9] = -13.0 / 18 * psi - 13.0 / 3 * t10 + t58 + (4.0 / 3) * t59 + (17.0 / 6) * t60 + t61 +
           (17.0 / 36) * t7 + (2.0 / 9) * t8 - 1;
    c[10] = (19.0 / 2) * a0 * psi + (19.0 / 12) * psi + t28 - 7.0 / 2 * t59 - 11.0 / 2 * t60 -
            t62 - 11.0 / 12 * t7 - 7.0 / 12 * t8 + 1.0 / 2;
    c[11] = -7.0 / 6 * psi - 7 * t10 + t58 + 3 * t59 + (7.0 / 2) * t60 + t62 + (7.0 / 12) * t7 +
            (1.0 / 2) * t8 + 1.0 / 3;
    c[12] = (11.0 / 6) * a0 * psi + (11.0 / 36) * psi - t57 - 5.0 / 6 * t59 - 5.0 / 6 * t60 -
            t61 - 5.0 / 36 * t7 - 5.0 / 36 * t8;
    c[13] = 0;
    c[14] = t42;
    c[15] = -t41;
    c[16] = 3.0 / 2 - t48;
    c[17] = t54 - 3;
    c[18] = 11.0 / 6 - t55;
    c[19] = t56;
    c[20] = 0;
    c[21] = (-149 * psi - t28 - t49 + t63 + t64 - 11) / (11 * psi + 6 * t7 + t8 + 6);
    c[22] = (149.0 / 6) * psi - t10 - t40 - t67;
    c[23] = (8 * a0 - psi * t70 + 3 * psi + 2 * t10 + t38 * t71 + 32 * t38 - t70 + t72 + 14) /
            (t68 + 2);
    c[24] = (-24 * a0 + 2 * psi * t34 * t35 * t36 * t66 + 4 * t34 * t35 * t36 * t66 -
             96 * t38 - t63 - t64 - t71 - t72 - 38) /
            (t68 + 4);
    c[25] = (72 * a0 - 4 * psi * t69 + 15 * psi + 18 * t10 + 288 * t38 + 72 * t46 - 12 * t69 +
             t72 + 78) /
            (6 * psi + 18);
    c[26] = -a0 - t39 - t67;
    c[27] = 0;

    const int n = 28;
    for (int i = 0; i < n; i++) c[i] /= h;
    if (right) {
        for (int i = 0; i < n; i++) c[i] *= -1;
        reverse(c, n);
    }
}

/* ---------------------------------------------------------------- E4u_1 */
void e4u1_floating(const double* alpha, double h, double* c, int right)
{
    const double a0 = alpha[0], a1 = alpha[1];
    c[0] = (6 * a0 - 11) / 6;
    c[1] = 3 - 4 * a0;
    c[2] = (12 * a0 - 3) / 2;
    c[3] = -(12 * a0 - 1) / 3;
    c[4] = a0;
    c[5] = (3 * a1 - 1) / 3;
    c[6] = -(8 * a1 + 1) / 2;
    c[7] = 6 * a1 + 1;
    c[8] = -(24 * a1 + 1) / 6;
    c[9] = a1;
    c[10] = -(168 * a1 + 54 * a0 - 11) / 138;
    c[11] = (112 * a1 + 36 * a0 - 15) / 23;
    c[12] = -(336 * a1 + 108 * a0 + 1) / 46;
    c[13] = (336 * a1 + 108 * a0 + 47) / 69;
    c[14] = -(28 * a1 + 9 * a0 + 2) / 23;
    const int n = 15;
    for (int i = 0; i < n; i++) c[i] /= h;
    if (right) {
        for (int i = 0; i < n; i++) c[i] *= -1;
        reverse(c, n);
    }
}

void e4u1_dirichlet(const double* alpha, double h, double* c, int right)
{
    const double a0 = alpha[0], a1 = alpha[1];
    c[0] = (3 * a1 - 1) / 3;
    c[1] = -(8 * a1 + 1) / 2;
    c[2] = 6 * a1 + 1;
    c[3] = -(24 * a1 + 1) / 6;
    c[4] = a1;
    c[5] = -(168 * a1 + 54 * a0 - 11) / 138;
    c[6] = (112 * a1 + 36 * a0 - 15) / 23;
    c[7] = -(336 * a1 + 108 * a0 + 1) / 46;
    c[8] = (336 * a1 + 108 * a0 + 47) / 69;
    c[9] = -(28 * a1 + 9 * a0 + 2) / 23;
    const int n = 10;
    for (int i = 0; i < n; i++) c[i] /= h;
    if (right) {
        for (int i = 0; i < n; i++) c[i] *= -1;
        reverse(c, n);
    }
}

/* ---------------------------------------------------------------- E4_2 */
void e42_floating(double h, double psi, double* c, int right)
{
    double t6 = -3 * psi;
    double t7 = psi * psi;
    double t9 = t7 * psi;
    double t11 = t7 * t7;
    double t12 = -t11;
    double t18 = 3 * psi;
    double t33 = -1 + psi;
    double t5 = 2 * psi;
    double t31 = 6 * psi;
    c[0] = t5;
    c[1] = (6 + t12 + t6 - 11 * t7 - 6 * t9) / 3.;
    c[2] = -5 + t11 + t6 + 6 * t7 + 5 * t9;
    c[3] = 4 + t12 + t18 - 3 * t7 - 4 * t9;
    c[4] = (-3 + t11 + t6 + 2 * t7 + 3 * t9) / 3.;
    c[5] = 12 / (6 + 5 * psi + t7);
    c[6] = -2 * psi;
    c[7] = -5 + t31;
    c[8] = (-2 * (4 + t18) * t33) / (2 + psi);
    c[9] = (t33 * (3 + t5)) / (3 + psi);
    c[10] = -psi / 12.;
    c[11] = (72 + t11 + t31 + 11 * t7 + 6 * t9) / 72.;
    c[12] = (-48 + t12 - 6 * t7 - 5 * t9) / 24.;
    c[13] = (24 + t11 + 3 * t7 + 4 * t9) / 24.;
    c[14] = -(t7 * (2 + t18 + t7)) / 72.;
    const int n = 15;
    for (int i = 0; i < n; i++) c[i] /= (h * h);
    if (right) reverse(c, n);
}

void e42_dirichlet(double h, double psi, double* c, int right)
{
    double t7 = psi * psi;
    double t9 = t7 * psi;
    double t11 = t7 * t7;
    double t12 = -t11;
    double t18 = 3 * psi;
    double t33 = -1 + psi;
    double t5 = 2 * psi;
    double t31 = 6 * psi;
    c[0] = 12 / (6 + 5 * psi + t7);
    c[1] = -2 * psi;
    c[2] = -5 + t31;
    c[3] = (-2 * (4 + t18) * t33) / (2 + psi);
    c[4] = (t33 * (3 + t5)) / (3 + psi);
    c[5] = -psi / 12.;
    c[6] = (72 + t11 + t31 + 11 * t7 + 6 * t9) / 72.;
    c[7] = (-48 + t12 - 6 * t7 - 5 * t9) / 24.;
    c[8] = (24 + t11 + 3 * t7 + 4 * t9) / 24.;
    c[9] = -(t7 * (2 + t18 + t7)) / 72.;
    const int n = 10;
    for (int i = 0; i < n; i++) c[i] /= (h * h);
    if (right) reverse(c, n);
}

void e42_neumann(double h, double psi, double* c, double* x, int right)
{
    double t7 = psi * psi;
    double t10 = 11 * psi;
    double t28 = -1 + psi;
    double t12 = t7 * psi;
    double t36 = t7 * t7;
    double t37 = -t36;
    double t51 = -6 * t7;
    double t8 = 3 * t7;
    x[0] = -3;
    x[1] = 3 * t28;
    x[2] = 0;
    c[0] = (-3 * (7 + 18 * psi + t8)) / (6 + t10 + t12 + 6 * t7);
    c[1] = 4 * psi;
    c[2] = (8 + t10 - 21 * t7) / (2 + 2 * psi);
    c[3] = (-1 - 8 * psi + 9 * t7) / (2 + psi);
    c[4] = (-5 * psi * t28) / (2. * (3 + psi));
    c[5] = psi;
    c[6] = (-21 - 9 * psi + 3 * t12 + t37 + 16 * t7) / 6.;
    c[7] = (8 + 12 * psi - 4 * t12 + t36 - 15 * t7) / 2.;
    c[8] = (-1 - 15 * psi + 5 * t12 + t37 + 12 * t7) / 2.;
    c[9] = (psi * (12 - 7 * psi + t12 + t51)) / 6.;
    c[10] = -psi / 12.;
    c[11] = (72 + 6 * psi + 6 * t12 + t36 + 11 * t7) / 72.;
    c[12] = (-48 - 5 * t12 + t37 + t51) / 24.;
    c[13] = (24 + 4 * t12 + t36 + t8) / 24.;
    c[14] = -(t7 * (2 + 3 * psi + t7)) / 72.;
    for (int i = 0; i < 15; i++) c[i] /= (h * h);
    for (int i = 0; i < 3; i++) x[i] /= h;
    if (right) {
        reverse(c, 15);
        reverse(x, 3);
        for (int i = 0; i < 3; i++) x[i] *= -1;
    }
}

/* ---------------------------------------------------------------- E2_2 */
void e22_floating(double h, double psi, double* c, int right)
{
    double t6 = psi * psi;
    double t8 = psi * psi * psi;
    double t9 = -1 * t8;
    double t5 = -2 * psi;
    c[0] = psi;
    c[1] = (2 + t5 + -3 * t6 + t9) * 0.5;
    c[2] = -2 + 2 * t6 + t8;
    c[3] = (2 + -1 * t6 + t9) * 0.5;
    c[4] = (2 + 4 * psi) * 1.0 / (2 + 3 * psi + t6);
    c[5] = t5;
    c[6] = (-2 + 4 * t6) * 1.0 / (1 + psi);
    c[7] = (2 + -2 * t6) * 1.0 / (2 + psi);
    for (int i = 0; i < 8; i++) c[i] /= (h * h);
    if (right) reverse(c, 8);
}

void e22_dirichlet(double h, double psi, double* c, int right)
{
    double t6 = psi * psi;
    double t5 = -2 * psi;
    c[0] = (2 + 4 * psi) * 1.0 / (2 + 3 * psi + t6);
    c[1] = t5;
    c[2] = (-2 + 4 * t6) * 1.0 / (1 + psi);
    c[3] = (2 + -2 * t6) * 1.0 / (2 + psi);
    for (int i = 0; i < 4; i++) c[i] /= (h * h);
    if (right) reverse(c, 4);
}

void e22_neumann(double h, double psi, double* c, double* x, int right)
{
    double t5 = 2 * psi;
    double t9 = psi * psi;
    double t18 = -1 + psi;
    double t23 = psi * psi * psi;
    x[0] = -2;
    x[1] = 2 * t18;
    c[0] = (4 + 8 * psi) * -1 * 1.0 / (2 + 3 * psi + t9);
    c[1] = t5;
    c[2] = (2 + t5 + -4 * t9) * 1.0 / (1 + psi);
    c[3] = 2 * psi * 1.0 / (2 + psi) * t18;
    c[4] = psi;
    c[5] = (-4 + -1 * t23 + t9) * 0.5;
    c[6] = 2 + t23 + -2 * t9;
    c[7] = (2 + -3 * psi + t9) * -0.5 * psi;
    for (int i = 0; i < 8; i++) c[i] /= (h * h);
    for (int i = 0; i < 2; i++) x[i] /= h;
    if (right) {
        reverse(c, 8);
        reverse(x, 2);
        for (int i = 0; i < 2; i++) x[i] *= -1;
    }
}


// Interpolants of the cut-point rows (E2_2.cpp:31-86, E4_2.cpp:31-183): weights
// over the boundary and the t - 1 points beyond it at offset y (in the
// i + y form of stencil.hpp:213-270), i the closest point's distance from
// the boundary point.
int e22_interp_wall(int i, double y, double psi, double* c, bool right)
{
    if (right) {
        const double t6 = 1 + psi, t7 = 1.0 / (t6), t8 = -1 + y, t9 = psi * t8, t5 = -1 + psi;
        const double t12 = -1 * psi * y, t20 = 1 + y;
        if (i == 0) {
            c[0] = (psi + y + psi * y) * t5;
            c[1] = 1 + t12 + -1 * psi * psi * t20 + y;
            c[2] = psi * t20;
        } else if (i == 1) {
            c[0] = (t9 + y) * t5 * t7;
            c[1] = psi + t12;
            c[2] = (1 + t9 + y) * t7;
        }
    } else {
        const double t8 = -1 + y, t11 = psi * y, t13 = -1 + psi, t17 = 1 + psi, t18 = 1.0 / (t17);
        if (i == 0) {
            c[0] = psi + -1 * psi * y;
            c[1] = 1 + t11 + psi * psi * t8 + -1 * y;
            c[2] = (psi * t8 + y) * -1 * t13;
        } else if (i == 1) {
            c[0] = (-1 + psi + t11 + y) * -1 * t18;
            c[1] = (1 + y) * psi;
            c[2] = (psi + t11 + y) * -1 * t13 * t18;
        }
    }
    return 3;
}

int e42_interp_wall(int i, double y, double psi, double* c, bool right)
{
    const double y2 = y * y, y3 = y2 * y;
    const double psi2 = psi * psi, psi3 = psi2 * psi, psi4 = psi2 * psi2;
    if (right) {
        if (i == 0) {
            c[0] = ((-1 + psi) * (6 * y * (2 + 3 * y + y2) + 6 * psi * (2 + 8 * y + 6 * y2 + y3) +
                                  psi3 * (6 + 11 * y + 6 * y2 + y3) + 2 * psi2 * (9 + 22 * y + 12 * y2 + 2 * y3))) /
                   36.;
            c[1] = (6 * y * (3 + 4 * y + y2) + 6 * psi * (3 + 8 * y + 3 * y2) - 3 * psi2 * (-2 + 5 * y + 6 * y2 + y3) -
                    psi4 * (6 + 11 * y + 6 * y2 + y3) - 2 * psi3 * (9 + 22 * y + 12 * y2 + 2 * y3)) /
                   12.;
            c[2] = (-6 * y * (6 + 5 * y + y2) - 6 * psi * (6 + 10 * y + 3 * y2) + 6 * psi2 * (1 + 8 * y + 6 * y2 + y3) +
                    psi4 * (6 + 11 * y + 6 * y2 + y3) + psi3 * (24 + 55 * y + 30 * y2 + 5 * y3)) /
                   12.;
            c[3] = (-6 * psi * (-5 - y + 3 * y2 + y3) - 6 * psi3 * (5 + 11 * y + 6 * y2 + y3) +
                    6 * (6 + 11 * y + 6 * y2 + y3) - psi4 * (6 + 11 * y + 6 * y2 + y3) -
                    psi2 * (30 + 103 * y + 66 * y2 + 11 * y3)) /
                   36.;
            c[4] = (psi * (6 + 11 * y + 6 * y2 + y3)) / 6.;
        } else if (i == 1) {
            c[0] = ((-1 + psi) * (psi * (-1 + y) + y) * (2 + 3 * y + y2)) / (2. * (3 + psi));
            c[1] = -((1 + y) * (2 * psi * (3 + y) - 2 * y * (3 + y) + 3 * psi2 * (-2 + y + y2))) / (2. * (2 + psi));
            c[2] = ((2 + y) * (psi * (3 + y) - y * (3 + y) + 3 * psi2 * (-1 + y2))) / (2. * (1 + psi));
            c[3] = (psi * (2 + y - 2 * y2 - y3)) / 2.;
            c[4] = ((3 + 3 * psi * (-1 + y) + y) * (2 + 3 * y + y2)) / (6 + 11 * psi + 6 * psi2 + psi3);
        } else {
            c[0] = (y * (-24 * (-1 + y2) + 2 * psi2 * (-2 - y + 2 * y2 + y3) + 3 * psi3 * (-2 - y + 2 * y2 + y3) +
                         psi4 * (-2 - y + 2 * y2 + y3))) /
                   144.;
            c[1] = (y * (24 * (-2 + y + y2) + psi4 * (2 + y - 2 * y2 - y3) - 3 * psi2 * (-2 - y + 2 * y2 + y3) -
                         4 * psi3 * (-2 - y + 2 * y2 + y3))) /
                   48.;
            c[2] = ((-24 + 6 * psi2 * y + 5 * psi3 * y + psi4 * y) * (-2 - y + 2 * y2 + y3)) / 48.;
            c[3] = -(y * (-24 * (2 + 3 * y + y2) + 6 * psi * (-2 - y + 2 * y2 + y3) +
                          11 * psi2 * (-2 - y + 2 * y2 + y3) + 6 * psi3 * (-2 - y + 2 * y2 + y3) +
                          psi4 * (-2 - y + 2 * y2 + y3))) /
                   144.;
            c[4] = (psi * y * (-2 - y + 2 * y2 + y3)) / 24.;
        }
    } else {
        if (i == 0) {
            c[0] = -(psi * (-6 + 11 * y - 6 * y2 + y3)) / 6.;
            c[1] = (-6 * (-6 + 11 * y - 6 * y2 + y3) + psi4 * (-6 + 11 * y - 6 * y2 + y3) +
                    6 * psi3 * (-5 + 11 * y - 6 * y2 + y3) + 6 * psi * (5 - y - 3 * y2 + y3) +
                    psi2 * (-30 + 103 * y - 66 * y2 + 11 * y3)) /
                   36.;
            c[2] = (6 * y * (6 - 5 * y + y2) - 6 * psi * (6 - 10 * y + 3 * y2) + psi3 * (24 - 55 * y + 30 * y2 - 5 * y3) -
                    6 * psi2 * (-1 + 8 * y - 6 * y2 + y3) - psi4 * (-6 + 11 * y - 6 * y2 + y3)) /
                   12.;
            c[3] = (-6 * y * (3 - 4 * y + y2) + 6 * psi * (3 - 8 * y + 3 * y2) + 3 * psi2 * (2 + 5 * y - 6 * y2 + y3) +
                    psi4 * (-6 + 11 * y - 6 * y2 + y3) + 2 * psi3 * (-9 + 22 * y - 12 * y2 + 2 * y3)) /
                   12.;
            c[4] = -((-1 + psi) * (6 * y * (2 - 3 * y + y2) + 6 * psi * (-2 + 8 * y - 6 * y2 + y3) +
                                   psi3 * (-6 + 11 * y - 6 * y2 + y3) + 2 * psi2 * (-9 + 22 * y - 12 * y2 + 2 * y3))) /
                   36.;
        } else if (i == 1) {
            c[0] = -(((2 - 3 * y + y2) * (-3 + y + 3 * psi * (1 + y))) / ((1 + psi) * (2 + psi) * (3 + psi)));
            c[1] = (psi * (2 - y - 2 * y2 + y3)) / 2.;
            c[2] = -((-2 + y) * (-(psi * (-3 + y)) - (-3 + y) * y + 3 * psi2 * (-1 + y2))) / (2. * (1 + psi));
            c[3] = ((-1 + y) * (-2 * psi * (-3 + y) - 2 * (-3 + y) * y + 3 * psi2 * (-2 - y + y2))) / (2. * (2 + psi));
            c[4] = -((-1 + psi) * (psi + y + psi * y) * (2 - 3 * y + y2)) / (2. * (3 + psi));
        } else {
            c[0] = (psi * y * (2 - y - 2 * y2 + y3)) / 24.;
            c[1] = -(y * (24 * (2 - 3 * y + y2) + 6 * psi * (2 - y - 2 * y2 + y3) + 11 * psi2 * (2 - y - 2 * y2 + y3) +
                          6 * psi3 * (2 - y - 2 * y2 + y3) + psi4 * (2 - y - 2 * y2 + y3))) /
                   144.;
            c[2] = ((24 + 6 * psi2 * y + 5 * psi3 * y + psi4 * y) * (2 - y - 2 * y2 + y3)) / 48.;
            c[3] = (y * (24 * (2 + y - y2) + psi4 * (-2 + y + 2 * y2 - y3) - 3 * psi2 * (2 - y - 2 * y2 + y3) -
                         4 * psi3 * (2 - y - 2 * y2 + y3))) /
                   48.;
            c[4] = (y * (24 * (-1 + y2) + 2 * psi2 * (2 - y - 2 * y2 + y3) + 3 * psi3 * (2 - y - 2 * y2 + y3) +
                         psi4 * (2 - y - 2 * y2 + y3))) /
                   144.;
        }
    }
    return 5;
}

/* ------------------------------------------------------------- E2-poly
 * polyE2_1.cpp:166-245 closures (fa: the 6 floating parameters, da: the 3
 * Dirichlet ones) and :52-138 wall interpolants (fa[0..3], ia: the 4
 * interpolant parameters). */
void e2p_closure_floating(const double* fa, double h, double psi, double* c, bool right)
{
    const double q = 1.0 / (1 + psi), s = 2 + psi;
    for (int row = 0; row < 3; ++row) {
        const double lo = fa[2 * row], hi = fa[2 * row + 1], him = -1 + hi;
        double* o = c + 4 * row;
        // (row 0 multiplies its factors in another order than rows 1 and 2)
        o[0] = row == 0 ? 0.5 * q * him : 0.5 * him * q;
        o[1] = (-1 + lo) * 0.5;
        o[2] = -1 * lo + (row == 0 ? -0.5 * s * q * him : -0.5 * s * him * q);
        o[3] = (row == 0 ? lo + hi : hi + lo) * 0.5;
    }
    for (int i = 0; i < 12; ++i) c[i] /= h;
    if (right) {
        for (int i = 0; i < 12; ++i) c[i] *= -1;
        reverse(c, 12);
    }
}

void e2p_closure_dirichlet(const double* da, double h, double psi, double* c, bool right)
{
    const double d0 = da[0], d1 = da[1], d2 = da[2];
    const double q = 1.0 / (1 + psi), d0m = -1 + d0, d2m = -1 + d2, s = 2 + psi;
    const double p2 = 2 * psi, d2x3 = 3 * d2, d2p2 = 2 * psi * d2;
    const double inv = 1.0 / (1 + p2 + d2x3 + d2p2);
    const double d1p2 = 2 * psi * d1, d1d0x3 = 3 * d1 * d0, d1d0p2 = 2 * psi * d1 * d0;
    c[0] = 0.5 * q * d0m;
    c[1] = (-1 + -2 * psi + d1 + d1p2 + d1d0x3 + d1d0p2 + -3 * d0 + -2 * psi * d0) * 0.5 * inv;
    c[2] = (-2 * psi * d1 + -3 * d2 + -2 * psi * d2 + -1 * d1 + 3 * d0 + 2 * psi * d0 + -3 * d1 * d0 +
            -2 * psi * d1 * d0) *
               inv +
           -0.5 * s * q * d0m;
    c[3] = (d1 + d1p2 + d1d0x3 + d1d0p2 + d2x3 + d2p2 + -2 * d0 + 3 * d2 * d0 + 2 * psi * d2 * d0) * 0.5 * inv;
    c[4] = 0.5 * d2m * q;
    c[5] = (-1 + d1) * 0.5;
    c[6] = -1 * d1 + -0.5 * s * d2m * q;
    c[7] = (d1 + d2) * 0.5;
    for (int i = 0; i < 8; ++i) c[i] /= h;
    if (right) {
        for (int i = 0; i < 8; ++i) c[i] *= -1;
        reverse(c, 8);
    }
}

int e2p_interp_wall(const double* fa, const double* ia, int i, double y, double psi, double* c, bool right)
{
    if (i != 0 && i != 1) return 4;
    const double q = 1.0 / (1 + psi), my = -1 * y;
    // row i uses the parameter pairs (fa, ia)[2 i], [2 i + 1]
    const double fl = fa[2 * i], fh = fa[2 * i + 1], il = ia[2 * i], ih = ia[2 * i + 1];
    const double fly = fl * y, fhy = fh * y;
    // the row of node 1 - psi carries psi (i = 0) or nothing (i = 1)
    const double lead = i == 0 ? 1 + psi : 1;
    if (right) {
        c[0] = (i == 0 ? fly + fhy + il + ih : ih + fly + fhy + il) * 0.5;
        c[1] = i == 0 ? (-1 * psi + my + (fly + il) * -2 +
                         (my + -2 * ih + -1 * psi * ih + -2 * fh * y + -1 * psi * fh * y) * q) *
                            0.5
                      : (my + (fly + il) * -2 +
                         (psi + -2 * ih + -1 * psi * ih + my + -2 * fh * y + -1 * psi * fh * y) * q) *
                            0.5;
        c[2] = (lead + fly + il + y) * 0.5;
        c[3] = (i == 0 ? 1 + psi + fhy + ih + y : 1 + ih + fhy + y) * 0.5 * q;
    } else {
        c[0] = (i == 0 ? 1 + psi + ih + my + fhy : 1 + fhy + ih + my) * 0.5 * q;
        c[1] = (lead + fly + il + my) * 0.5;
        c[2] = i == 0 ? (-1 * psi + (fly + il) * -2 + y +
                         (-2 * ih + -1 * psi * ih + y + -2 * fh * y + -1 * psi * fh * y) * q) *
                            0.5
                      : ((fly + il) * -2 + y +
                         (psi + -2 * ih + -1 * psi * ih + y + -2 * fh * y + -1 * psi * fh * y) * q) *
                            0.5;
        c[3] = (i == 0 ? ih + fly + il + fhy : fhy + ih + fly + il) * 0.5;
    }
    return 4;
}

}  // namespace

info stencil::query_max() const
{
    switch (kind_) {
    case scheme::E2_2: return {1, 2, 4, 2};
    case scheme::E4_2: return {2, 3, 5, 3};
    case scheme::E4_1: return {2, 5, 7, 0};
    case scheme::E4u_1: return {2, 3, 5, 0};
    case scheme::E2poly_1: return {1, 3, 4, 0};
    }
    return {};
}

info stencil::query(bcs::type b) const
{
    info m = query_max();
    switch (b) {
    case bcs::type::Dirichlet: m.r -= 1; m.nextra = 0; return m;
    case bcs::type::Floating: m.nextra = 0; return m;
    case bcs::type::Neumann:
        if (kind_ == scheme::E4_1 || kind_ == scheme::E4u_1 || kind_ == scheme::E2poly_1) return {};
        return m;
    }
    return {};
}

void stencil::interior(double h, double* c) const
{
    switch (kind_) {
    case scheme::E2_2:
        c[0] = 1 / (h * h);
        c[1] = -2 / (h * h);
        c[2] = 1 / (h * h);
        break;
    case scheme::E4_2:
        c[0] = -1 / (12. * (h * h));
        c[1] = 4 / (3. * (h * h));
        c[2] = -5 / (2. * (h * h));
        c[3] = 4 / (3. * (h * h));
        c[4] = -1 / (12. * (h * h));
        break;
    case scheme::E4_1:
        c[0] = 0.083333333333333329;
        c[1] = -0.66666666666666663;
        c[2] = 0;
        c[3] = 0.66666666666666663;
        c[4] = -0.083333333333333329;
        for (int i = 0; i < 5; ++i) c[i] /= h;
        break;
    case scheme::E4u_1:
        c[0] = 1 / (12 * h);
        c[1] = -2 / (3 * h);
        c[2] = 0;
        c[3] = -c[1];
        c[4] = -c[0];
        break;
    case scheme::E2poly_1:
        c[0] = -0.5 / h;
        c[1] = 0 / h;
        c[2] = 0.5 / h;
        break;
    }
}

void stencil::nbs(double h, bcs::type b, double psi, bool right, double* c, double* extra) const
{
    const double* a = alpha_.data();
    switch (kind_) {
    case scheme::E4_1:
        if (b == bcs::type::Floating) e41_floating(a, h, psi, c, right);
        else if (b == bcs::type::Dirichlet) e41_dirichlet(a, h, psi, c, right);
        break;
    case scheme::E4u_1:
        if (b == bcs::type::Floating) e4u1_floating(a, h, c, right);
        else if (b == bcs::type::Dirichlet) e4u1_dirichlet(a, h, c, right);
        break;
    case scheme::E4_2:
        if (b == bcs::type::Floating) e42_floating(h, psi, c, right);
        else if (b == bcs::type::Dirichlet) e42_dirichlet(h, psi, c, right);
        else e42_neumann(h, psi, c, extra, right);
        break;
    case scheme::E2_2:
        if (b == bcs::type::Floating) e22_floating(h, psi, c, right);
        else if (b == bcs::type::Dirichlet) e22_dirichlet(h, psi, c, right);
        else e22_neumann(h, psi, c, extra, right);
        break;
    case scheme::E2poly_1:
        if (b == bcs::type::Floating) e2p_closure_floating(a, h, psi, c, right);
        else if (b == bcs::type::Dirichlet) e2p_closure_dirichlet(a + 6, h, psi, c, right);
        break;
    }
}

std::optional<stencil> stencil::make(int kind, const double* alpha, int nalpha)
{
    std::array<double, kMaxAlpha> a{};
    for (int i = 0; i < nalpha && i < kMaxAlpha; ++i) a[i] = alpha[i];
    if (kind < 0 || kind > 4) return std::nullopt;
    // E4_1 denominator is singular for alpha[1] < 197/288 (src/stencils/E4_1.cpp:34-41)
    if (kind == (int)scheme::E4_1 && a[1] < 197.0 / 288.0) return std::nullopt;
    return stencil{(scheme)kind, a};
}

// stencil.hpp:29 query_interp: E2 2 interior points over 3 wall points, E4 5
// and 5; the first-derivative schemes have none
interp_info stencil::query_interp() const
{
    if (kind_ == scheme::E2_2) return {2, 3};
    if (kind_ == scheme::E4_2) return {5, 5};
    if (kind_ == scheme::E2poly_1) return {2, 4};
    return {};
}

int stencil::interp_interior(double y, double* c, int cap) const
{
    if (kind_ == scheme::E2_2 || kind_ == scheme::E2poly_1) {
        if (y > 0) {
            c[0] = 1 + -1 * y;
            c[1] = y;
        } else {
            c[0] = -1 * y;
            c[1] = 1 + y;
        }
        return 2;
    }
    if (kind_ == scheme::E4_2) {
        const double y2 = y * y, y3 = y * y * y, y4 = y2 * y2;
        c[0] = (y * (2 - y - 2 * y2 + y3)) / 24.;
        c[1] = (y * (-4 + 4 * y + y2 - y3)) / 6.;
        c[2] = (4 - 5 * y2 + y4) / 4.;
        c[3] = -(y * (-4 - 4 * y + y2 + y3)) / 6.;
        c[4] = (y * (-2 - y + 2 * y2 + y3)) / 24.;
        return 5;
    }
    return cap;
}

int stencil::interp_wall(int i, double y, double psi, double* c, bool right, int cap) const
{
    if (kind_ == scheme::E2_2) return e22_interp_wall(i, y, psi, c, right);
    if (kind_ == scheme::E4_2) return e42_interp_wall(i, y, psi, c, right);
    if (kind_ == scheme::E2poly_1) return e2p_interp_wall(alpha_.data(), alpha_.data() + 9, i, y, psi, c, right);
    return cap;
}

}  // namespace ccs::stencils
