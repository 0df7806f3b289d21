/*
 * oracle.c -- FP64 CPU restatement of PhySim's Newton inner loop (tet FEM).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for libpgn and the timed CPU
 * baseline ("port").  Never linked into, or called by, the product path.
 *
 * Element loops use OpenMP like Simulable.cpp's `#pragma omp parallel for`
 * (Simulable.cpp:1038-1047,1099-1108,1160-1169); assembly and the energy sum
 * are serial in element order like Simulable.cpp:333-334,396-397,580-581.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ======================================================================= */
/* Mesh generator                                                          */
/* ======================================================================= */

int orc_tet_grid_counts(const int dims[3], int* num_nodes, int* num_tets) {
  *num_nodes = (dims[0] + 1) * (dims[1] + 1) * (dims[2] + 1);
  *num_tets = 5 * dims[0] * dims[1] * dims[2];
  return *num_tets;
}

/* MeshUtils.cpp:927-991: nodes x-fastest on a regular grid, 5 tets per hex
 * cell with the same split in every cell. */
int orc_tet_grid(const double scale[3], const int dims[3], double* V, int* T) {
  const int nx = dims[0] + 1, ny = dims[1] + 1, nz = dims[2] + 1;
  const double h[3] = {scale[0] / dims[0], scale[1] / dims[1],
                       scale[2] / dims[2]};
  int c = 0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        V[3 * c + 0] = i * h[0];
        V[3 * c + 1] = j * h[1];
        V[3 * c + 2] = k * h[2];
        ++c;
      }
  const int line = nx, slab = nx * ny;
  int e = 0;
  for (int k = 0; k < dims[2]; ++k)
    for (int j = 0; j < dims[1]; ++j)
      for (int i = 0; i < dims[0]; ++i) {
        const int o = slab * k + line * j + i;
        /* hex corners: bit0 = +x, bit1 = +y, bit2 = +z */
        const int h0 = o, h1 = o + 1, h2 = o + line, h3 = o + line + 1;
        const int h4 = o + slab, h5 = o + slab + 1, h6 = o + slab + line,
                  h7 = o + slab + line + 1;
        const int tets[5][4] = {{h0, h1, h2, h4},
                                {h6, h7, h4, h2},
                                {h5, h4, h7, h1},
                                {h3, h2, h1, h7},
                                {h1, h2, h4, h7}};
        for (int t = 0; t < 5; ++t) {
          for (int a = 0; a < 4; ++a) T[4 * e + a] = tets[t][a];
          ++e;
        }
      }
  return e;
}

void orc_lame_from_young_poisson(double young, double poisson, double* lame1,
                                 double* lame2) {
  const double t0 = 1.0 + poisson, t1 = 1.0 - 2.0 * poisson;
  *lame1 = (young * poisson) / (t0 * t1);
  *lame2 = young / (2.0 * t0);
}

/* ======================================================================= */
/* Material models.  Column-major vec(F): Fv[i + 3 j] = F(i, j).            */
/* Closed forms derived independently of the reference's Maple output:     */
/*   CoNH: Psi = mu/2 (I1 - 3) - mu ln J + lambda/2 ln^2 J                  */
/*         P   = mu (F - F^-T) + lambda ln J F^-T                           */
/*   StVK: Psi = lambda/2 tr(E)^2 + mu tr(E^2), E = (F^T F - I)/2           */
/*         P   = F S, S = 2 mu E + lambda tr(E) I                           */
/* (MaterialModels.h:166-299 selects between them; lame1 = lambda,          */
/*  lame2 = mu.)                                                            */
/* ======================================================================= */

#define FV(i, j) Fv[(i) + 3 * (j)]

static double det3(const double* Fv) {
  return FV(0, 0) * (FV(1, 1) * FV(2, 2) - FV(1, 2) * FV(2, 1)) -
         FV(0, 1) * (FV(1, 0) * FV(2, 2) - FV(1, 2) * FV(2, 0)) +
         FV(0, 2) * (FV(1, 0) * FV(2, 1) - FV(1, 1) * FV(2, 0));
}

/* inverse (column-major in and out); returns det */
static double inv3(const double* Fv, double* Iv) {
  const double d = det3(Fv);
  const double id = 1.0 / d;
#define IV(i, j) Iv[(i) + 3 * (j)]
  IV(0, 0) = (FV(1, 1) * FV(2, 2) - FV(1, 2) * FV(2, 1)) * id;
  IV(0, 1) = (FV(0, 2) * FV(2, 1) - FV(0, 1) * FV(2, 2)) * id;
  IV(0, 2) = (FV(0, 1) * FV(1, 2) - FV(0, 2) * FV(1, 1)) * id;
  IV(1, 0) = (FV(1, 2) * FV(2, 0) - FV(1, 0) * FV(2, 2)) * id;
  IV(1, 1) = (FV(0, 0) * FV(2, 2) - FV(0, 2) * FV(2, 0)) * id;
  IV(1, 2) = (FV(0, 2) * FV(1, 0) - FV(0, 0) * FV(1, 2)) * id;
  IV(2, 0) = (FV(1, 0) * FV(2, 1) - FV(1, 1) * FV(2, 0)) * id;
  IV(2, 1) = (FV(0, 1) * FV(2, 0) - FV(0, 0) * FV(2, 1)) * id;
  IV(2, 2) = (FV(0, 0) * FV(1, 1) - FV(0, 1) * FV(1, 0)) * id;
#undef IV
  return d;
}

static void stvk_S(const double* Fv, double lame1, double lame2, double S[9]) {
  double E[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double c = 0;
      for (int k = 0; k < 3; ++k) c += FV(k, i) * FV(k, j);
      E[i + 3 * j] = 0.5 * (c - (i == j ? 1.0 : 0.0));
    }
  const double tr = E[0] + E[4] + E[8];
  for (int q = 0; q < 9; ++q) S[q] = 2.0 * lame2 * E[q];
  S[0] += lame1 * tr;
  S[4] += lame1 * tr;
  S[8] += lame1 * tr;
}

double orc_psi(int kind, const double* Fv, double lame1, double lame2) {
  if (kind == ORC_CONH) {
    double I1 = 0;
    for (int q = 0; q < 9; ++q) I1 += Fv[q] * Fv[q];
    const double J = det3(Fv);
    /* ln(J^2)/2 = ln|J|, as the reference's Maple energy uses log(J^2) */
    const double lnJ = 0.5 * log(J * J);
    return 0.5 * lame2 * (I1 - 3.0) - lame2 * lnJ + 0.5 * lame1 * lnJ * lnJ;
  } else {
    double E[9];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double c = 0;
        for (int k = 0; k < 3; ++k) c += FV(k, i) * FV(k, j);
        E[i + 3 * j] = 0.5 * (c - (i == j ? 1.0 : 0.0));
      }
    const double tr = E[0] + E[4] + E[8];
    double trE2 = 0;
    for (int q = 0; q < 9; ++q) trE2 += E[q] * E[q];
    return 0.5 * lame1 * tr * tr + lame2 * trE2;
  }
}

void orc_pk1(int kind, const double* Fv, double lame1, double lame2,
             double* P) {
  if (kind == ORC_CONH) {
    double Fi[9];
    const double J = inv3(Fv, Fi);
    const double lnJ = 0.5 * log(J * J);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        const double FiT = Fi[j + 3 * i]; /* F^-T (i,j) = F^-1 (j,i) */
        P[i + 3 * j] = lame2 * (FV(i, j) - FiT) + lame1 * lnJ * FiT;
      }
  } else {
    double S[9];
    stvk_S(Fv, lame1, lame2, S);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double c = 0;
        for (int k = 0; k < 3; ++k) c += FV(i, k) * S[k + 3 * j];
        P[i + 3 * j] = c;
      }
  }
}

/* C[p * 9 + q] = d P_p / d F_q, p = i + 3 j, q = m + 3 n */
void orc_dpdf(int kind, const double* Fv, double lame1, double lame2,
              double* C) {
  if (kind == ORC_CONH) {
    double Fi[9];
    const double J = inv3(Fv, Fi);
    const double lnJ = 0.5 * log(J * J);
    const double a = lame2 - lame1 * lnJ;
#define FI(i, j) Fi[(i) + 3 * (j)]
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i)
        for (int n = 0; n < 3; ++n)
          for (int m = 0; m < 3; ++m) {
            double v = a * FI(j, m) * FI(n, i) + lame1 * FI(n, m) * FI(j, i);
            if (i == m && j == n) v += lame2;
            C[(i + 3 * j) * 9 + (m + 3 * n)] = v;
          }
#undef FI
  } else {
    double S[9], FFt[9];
    stvk_S(Fv, lame1, lame2, S);
    for (int i = 0; i < 3; ++i)
      for (int m = 0; m < 3; ++m) {
        double c = 0;
        for (int k = 0; k < 3; ++k) c += FV(i, k) * FV(m, k);
        FFt[i + 3 * m] = c;
      }
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i)
        for (int n = 0; n < 3; ++n)
          for (int m = 0; m < 3; ++m) {
            double v = lame2 * FV(i, n) * FV(m, j) + lame1 * FV(m, n) * FV(i, j);
            if (i == m) v += S[n + 3 * j];
            if (j == n) v += lame2 * FFt[i + 3 * m];
            C[(i + 3 * j) * 9 + (m + 3 * n)] = v;
          }
  }
}
#undef FV

/* ======================================================================= */
/* Tet4 element                                                             */
/* ======================================================================= */

/* ShapeFunctions.cpp:89-133: J = N0 * B (3x3), BH0 = B * J^-1 (4x3),
 * DFDx(:, d) = vec_col(E_d * BH0) with E_d the 3x4 unit matrix whose
 * column-major entry d is 1 (node d/3, coordinate d%3).
 * EnergyElement_FEM.cpp:90: weight = (1/6) det(J) (Quadratures.h:134). */
void orc_tet_rest(const double X[12], double DFDx[108], double* V0) {
  static const double B[4][3] = {
      {-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  double Jm[9]; /* column-major */
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double c = 0;
      for (int a = 0; a < 4; ++a) c += X[3 * a + i] * B[a][j];
      Jm[i + 3 * j] = c;
    }
  double Ji[9];
  const double detJ = inv3(Jm, Ji);
  double BH[4][3];
  for (int a = 0; a < 4; ++a)
    for (int j = 0; j < 3; ++j) {
      double c = 0;
      for (int k = 0; k < 3; ++k) c += B[a][k] * Ji[k + 3 * j];
      BH[a][j] = c;
    }
  /* (E_d * BH)(i, j) = delta(i, d%3) * BH[d/3][j]; vec index r = i + 3 j */
  for (int r = 0; r < 9; ++r)
    for (int d = 0; d < 12; ++d) {
      const int i = r % 3, j = r / 3;
      DFDx[r * 12 + d] = (i == d % 3) ? BH[d / 3][j] : 0.0;
    }
  *V0 = (1.0 / 6.0) * detJ;
}

/* EnergyElement_FEM.cpp:125-222 with ShapeFunctions.cpp:135-158 (F). */
int orc_tet_egh(const double x[12], const double DFDx[108], double V0,
                int kind, double lame1, double lame2, double* E, double g[12],
                double H[144]) {
  double Fv[9];
  for (int r = 0; r < 9; ++r) {
    double c = 0;
    for (int d = 0; d < 12; ++d) c += DFDx[r * 12 + d] * x[d];
    Fv[r] = c;
  }
  /* Cell_Tetra.cpp:27-36 VolumeSpace at the current positions */
  double Ds[9];
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < 3; ++i) Ds[i + 3 * a] = x[3 * (a + 1) + i] - x[i];
  const int inverted = (det3(Ds) / 6.0) < 0;
  if (E) *E = inverted ? HUGE_VAL : V0 * orc_psi(kind, Fv, lame1, lame2);
  if (g) {
    double P[9];
    orc_pk1(kind, Fv, lame1, lame2, P);
    for (int d = 0; d < 12; ++d) {
      double c = 0;
      for (int r = 0; r < 9; ++r) c += DFDx[r * 12 + d] * P[r];
      g[d] = V0 * c;
    }
  }
  if (H) {
    double C[81], CD[9 * 12];
    orc_dpdf(kind, Fv, lame1, lame2, C);
    for (int p = 0; p < 9; ++p)
      for (int d = 0; d < 12; ++d) {
        double c = 0;
        for (int q = 0; q < 9; ++q) c += C[p * 9 + q] * DFDx[q * 12 + d];
        CD[p * 12 + d] = c;
      }
    for (int d = 0; d < 12; ++d)
      for (int e = 0; e < 12; ++e) {
        double c = 0;
        for (int p = 0; p < 9; ++p) c += DFDx[p * 12 + d] * CD[p * 12 + e];
        H[d * 12 + e] = V0 * c;
      }
  }
  return inverted;
}

/* ======================================================================= */
/* Model                                                                    */
/* ======================================================================= */

struct orc_model {
  int n, m;
  double* X0;
  int* T;
  unsigned char* kind;
  double *lame1, *lame2, *rho;
  double* DFDx; /* 108 per tet */
  double* V0;
  double* elem_mass; /* lumped mass per node of the element */
  double* sp;        /* 8 per element: membrane Dm^-1 (4), A0 t; hinge kB, theta0, shape0 */
  unsigned char* fixed; /* 3n */
  double* x;
  /* gravity BC */
  int has_g, g_stages, g_cur;
  double g_ini[3], g_end[3];
  /* force BC */
  int has_f, f_stages, f_cur;
  double *f_ini, *f_end;
  /* per-element outputs */
  double *eE, *eg, *eH;
  double escale; /* |elastic| + |load| energy of the last evaluation */
  /* lower pattern */
  long nnz;
  long* rowptr;
  int* col;
  long* slot; /* per element 144 local entries -> CSR slot or -1 */
};

/* nodes of an element that carry mass and load: tets 4, membranes 3 (their
 * fourth connectivity slot repeats node 2), hinges none */
static int mass_nodes(int kind) { return kind <= 1 ? 4 : (kind == 2 ? 3 : 0); }

orc_model* orc_model_create(int n_nodes, const double* X0, int n_tets,
                            const int* T, const unsigned char* kind,
                            const double* lame1, const double* lame2,
                            const double* rho) {
  return orc_model_create_mixed(n_nodes, X0, n_tets, T, kind, lame1, lame2, rho, NULL);
}

orc_model* orc_model_create_mixed(int n_nodes, const double* X0, int n_tets,
                                  const int* T, const unsigned char* kind,
                                  const double* lame1, const double* lame2,
                                  const double* rho, const double* extra) {
  orc_model* M = (orc_model*)calloc(1, sizeof(orc_model));
  M->n = n_nodes;
  M->m = n_tets;
  M->X0 = (double*)malloc(sizeof(double) * 3 * n_nodes);
  memcpy(M->X0, X0, sizeof(double) * 3 * n_nodes);
  M->x = (double*)malloc(sizeof(double) * 3 * n_nodes);
  memcpy(M->x, X0, sizeof(double) * 3 * n_nodes);
  M->T = (int*)malloc(sizeof(int) * 4 * n_tets);
  memcpy(M->T, T, sizeof(int) * 4 * n_tets);
  M->kind = (unsigned char*)malloc(n_tets);
  memcpy(M->kind, kind, n_tets);
  M->lame1 = (double*)malloc(sizeof(double) * n_tets);
  M->lame2 = (double*)malloc(sizeof(double) * n_tets);
  M->rho = (double*)malloc(sizeof(double) * n_tets);
  memcpy(M->lame1, lame1, sizeof(double) * n_tets);
  memcpy(M->lame2, lame2, sizeof(double) * n_tets);
  memcpy(M->rho, rho, sizeof(double) * n_tets);
  M->DFDx = (double*)malloc(sizeof(double) * 108 * n_tets);
  M->V0 = (double*)malloc(sizeof(double) * n_tets);
  M->elem_mass = (double*)malloc(sizeof(double) * n_tets);
  M->sp = (double*)calloc(8 * (size_t)(n_tets > 0 ? n_tets : 1), sizeof(double));
  M->fixed = (unsigned char*)calloc(3 * n_nodes, 1);
  M->eE = (double*)malloc(sizeof(double) * n_tets);
  M->eg = (double*)malloc(sizeof(double) * 12 * n_tets);
  M->eH = (double*)malloc(sizeof(double) * 144 * n_tets);
  /* EnergyElement_FEM::Init over elements (Simulable.cpp:963 omp loop) */
#pragma omp parallel for schedule(static)
  for (int e = 0; e < n_tets; ++e) {
    double Xe[12];
    for (int a = 0; a < 4; ++a)
      for (int i = 0; i < 3; ++i) Xe[3 * a + i] = X0[3 * T[4 * e + a] + i];
    if (kind[e] <= 1) {
      orc_tet_rest(Xe, M->DFDx + 108 * e, M->V0 + e);
      /* MassElement_Lumped.cpp:47-54: rho * VolumeBasis(Position_0) / nodes */
      M->elem_mass[e] = rho[e] * M->V0[e] / 4.0;
    } else if (kind[e] == 2) {
      /* membrane: extra[e] = thickness; VolumeBasis of a triangle is its
       * area (Face_Tri.cpp:33-39), with no thickness factor */
      double A;
      orc_membrane_rest(Xe, M->sp + 8 * e, &A);
      M->sp[8 * e + 4] = A * extra[e];
      M->V0[e] = A;
      M->elem_mass[e] = rho[e] * A / 3.0;
    } else {
      /* hinge: EnergyElement_DiscreteShells::Init (:50-54); extra[e] = kB */
      M->sp[8 * e] = extra[e];
      M->sp[8 * e + 1] = orc_hinge_theta(Xe);
      M->sp[8 * e + 2] = orc_hinge_shape(Xe);
      M->V0[e] = 0.0;
      M->elem_mass[e] = 0.0;
    }
  }
  /* Lower pattern as AMatrixSd stores it: node block (a, b) with a > b is
   * stored whole below the diagonal; node block (a, a) is stored whole
   * (BasicTypes.cpp:18-33, DerivativeAssembler.cpp:220). */
  const int N = 3 * n_nodes;
  /* node adjacency (lower) */
  int* deg = (int*)calloc(n_nodes, sizeof(int));
  for (int e = 0; e < n_tets; ++e)
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) {
        const int na = T[4 * e + a], nb = T[4 * e + b];
        if (na >= nb) deg[na]++;
      }
  long* aptr = (long*)malloc(sizeof(long) * (n_nodes + 1));
  aptr[0] = 0;
  for (int i = 0; i < n_nodes; ++i) aptr[i + 1] = aptr[i] + deg[i];
  int* adj = (int*)malloc(sizeof(int) * aptr[n_nodes]);
  memset(deg, 0, sizeof(int) * n_nodes);
  for (int e = 0; e < n_tets; ++e)
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) {
        const int na = T[4 * e + a], nb = T[4 * e + b];
        if (na >= nb) adj[aptr[na] + deg[na]++] = nb;
      }
  /* sort + unique per node */
  int* ucount = (int*)calloc(n_nodes, sizeof(int));
  for (int i = 0; i < n_nodes; ++i) {
    int* L = adj + aptr[i];
    const int c = deg[i];
    for (int p = 1; p < c; ++p) { /* insertion sort; lists are short */
      int v = L[p], q = p - 1;
      while (q >= 0 && L[q] > v) {
        L[q + 1] = L[q];
        --q;
      }
      L[q + 1] = v;
    }
    int u = 0;
    for (int p = 0; p < c; ++p)
      if (u == 0 || L[u - 1] != L[p]) L[u++] = L[p];
    ucount[i] = u;
  }
  M->rowptr = (long*)malloc(sizeof(long) * (N + 1));
  M->rowptr[0] = 0;
  for (int i = 0; i < n_nodes; ++i)
    for (int di = 0; di < 3; ++di)
      M->rowptr[3 * i + di + 1] = M->rowptr[3 * i + di] + 3L * ucount[i];
  M->nnz = M->rowptr[N];
  M->col = (int*)malloc(sizeof(int) * M->nnz);
  for (int i = 0; i < n_nodes; ++i)
    for (int di = 0; di < 3; ++di) {
      long p = M->rowptr[3 * i + di];
      for (int q = 0; q < ucount[i]; ++q)
        for (int dj = 0; dj < 3; ++dj) M->col[p++] = 3 * adj[aptr[i] + q] + dj;
    }
  /* element slot map: local (d, e) -> CSR index of the stored entry it is
   * added to, or -1 when the entry is the mirror of a stored one */
  M->slot = (long*)malloc(sizeof(long) * 144 * (long)n_tets);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < n_tets; ++e) {
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) {
        const int na = M->T[4 * e + a], nb = M->T[4 * e + b];
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) {
            const int d = 3 * a + i, f = 3 * b + j;
            long s = -1;
            /* The reference walks local blocks (a, b) with b <= a and adds
             * block (a, b) at (na, nb) when na >= nb, else its transpose at
             * (nb, na); local blocks with b > a are never read. */
            if (b <= a) {
              const int r = (na >= nb) ? 3 * na + i : 3 * nb + j;
              const int c = (na >= nb) ? 3 * nb + j : 3 * na + i;
              /* locate (r, c) */
              long lo = M->rowptr[r], hi = M->rowptr[r + 1] - 1;
              while (lo <= hi) {
                long mid = (lo + hi) / 2;
                if (M->col[mid] == c) {
                  s = mid;
                  break;
                }
                if (M->col[mid] < c) lo = mid + 1;
                else hi = mid - 1;
              }
            }
            M->slot[144L * e + d * 12 + f] = s;
          }
      }
  }
  free(deg);
  free(aptr);
  free(adj);
  free(ucount);
  return M;
}

void orc_model_destroy(orc_model* M) {
  if (!M) return;
  free(M->X0);
  free(M->x);
  free(M->T);
  free(M->kind);
  free(M->lame1);
  free(M->lame2);
  free(M->rho);
  free(M->DFDx);
  free(M->V0);
  free(M->elem_mass);
  free(M->sp);
  free(M->fixed);
  free(M->f_ini);
  free(M->f_end);
  free(M->eE);
  free(M->eg);
  free(M->eH);
  free(M->rowptr);
  free(M->col);
  free(M->slot);
  free(M);
}

int orc_model_num_dofs(const orc_model* M) { return 3 * M->n; }

void orc_model_set_fixed(orc_model* M, const unsigned char* mask) {
  memcpy(M->fixed, mask, 3 * M->n);
}

void orc_model_set_gravity(orc_model* M, const double g_ini[3],
                           const double g_end[3], int stages) {
  M->has_g = 1;
  M->g_stages = stages < 1 ? 1 : stages;
  M->g_cur = 0;
  memcpy(M->g_ini, g_ini, sizeof(double) * 3);
  memcpy(M->g_end, g_end, sizeof(double) * 3);
}

void orc_model_set_forces(orc_model* M, const double* f_ini,
                          const double* f_end, int stages) {
  const int N = 3 * M->n;
  M->has_f = 1;
  M->f_stages = stages < 1 ? 1 : stages;
  M->f_cur = 0;
  free(M->f_ini);
  free(M->f_end);
  M->f_ini = (double*)malloc(sizeof(double) * N);
  M->f_end = (double*)malloc(sizeof(double) * N);
  memcpy(M->f_ini, f_ini, sizeof(double) * N);
  memcpy(M->f_end, f_end, sizeof(double) * N);
}

/* BCondition.cpp:42 StepLoading; BCondition.h:128 IsLoaded */
int orc_model_step_load(orc_model* M) {
  int stepped = 0;
  if (M->has_g && M->g_cur < M->g_stages) {
    M->g_cur++;
    stepped = 1;
  }
  if (M->has_f && M->f_cur < M->f_stages) {
    M->f_cur++;
    stepped = 1;
  }
  return stepped;
}

int orc_model_fully_loaded(const orc_model* M) {
  if (M->has_g && M->g_cur != M->g_stages) return 0;
  if (M->has_f && M->f_cur != M->f_stages) return 0;
  return 1;
}

void orc_model_set_x(orc_model* M, const double* x) {
  memcpy(M->x, x, sizeof(double) * 3 * M->n);
}
void orc_model_get_x(const orc_model* M, double* x) {
  memcpy(x, M->x, sizeof(double) * 3 * M->n);
}

void orc_model_get_mass(const orc_model* M, double* mass3n) {
  memset(mass3n, 0, sizeof(double) * 3 * M->n);
  for (int e = 0; e < M->m; ++e)
    for (int a = 0; a < mass_nodes(M->kind[e]); ++a)
      for (int i = 0; i < 3; ++i) mass3n[3 * M->T[4 * e + a] + i] += M->elem_mass[e];
}

/* BCondition.h:140-155 GetCurValues with VectorAnimator_LerpValues */
static void cur_gravity(const orc_model* M, double g[3]) {
  if (!M->has_g) {
    g[0] = g[1] = g[2] = 0;
    return;
  }
  const double a = (double)M->g_cur / (double)M->g_stages;
  for (int i = 0; i < 3; ++i) g[i] = M->g_ini[i] + (M->g_end[i] - M->g_ini[i]) * a;
}

static double cur_force(const orc_model* M, int k) {
  const double a = (double)M->f_cur / (double)M->f_stages;
  return M->f_ini[k] + (M->f_end[k] - M->f_ini[k]) * a;
}

static void gather_x(const orc_model* M, int e, double xe[12]) {
  for (int a = 0; a < 4; ++a)
    for (int i = 0; i < 3; ++i) xe[3 * a + i] = M->x[3 * M->T[4 * e + a] + i];
}

/* IEnergyElement::ComputeAndStore_* of element e: 12-vector / 12x12 over its
 * four connectivity slots (a membrane's fourth slot gets zeros) */
static void elem_egh(const orc_model* M, int e, const double xe[12], double* E,
                     double* g, double* H) {
  const int k = M->kind[e];
  if (k <= 1) {
    orc_tet_egh(xe, M->DFDx + 108 * e, M->V0[e], k, M->lame1[e], M->lame2[e], E, g, H);
    return;
  }
  if (k == 3) {
    const double* p = M->sp + 8 * e;
    const double en = orc_hinge_egh(xe, p[0], p[1], p[2], g, H);
    if (E) *E = en;
    return;
  }
  const double* p = M->sp + 8 * e;
  double g9[9], H81[81];
  const double en = orc_membrane_egh(xe, p, p[4], M->lame1[e], M->lame2[e], g ? g9 : NULL,
                                     H ? H81 : NULL);
  if (E) *E = en;
  if (g) {
    memset(g, 0, 12 * sizeof(double));
    memcpy(g, g9, 9 * sizeof(double));
  }
  if (H) {
    memset(H, 0, 144 * sizeof(double));
    for (int i = 0; i < 9; ++i)
      for (int j = 0; j < 9; ++j) H[12 * i + j] = H81[9 * i + j];
  }
}

double orc_model_energy(orc_model* M) {
#pragma omp parallel for schedule(static)
  for (int e = 0; e < M->m; ++e) {
    double xe[12];
    gather_x(M, e, xe);
    elem_egh(M, e, xe, M->eE + e, NULL, NULL);
  }
  double E = 0;
  for (int e = 0; e < M->m; ++e) E += M->eE[e];
  const double Eel = E;
  /* MassElement_Lumped.cpp:56-68: E = x . (-m g) per mass element */
  double g[3];
  cur_gravity(M, g);
  double Eload = 0;
  for (int e = 0; e < M->m; ++e) {
    double Ee = 0;
    for (int a = 0; a < mass_nodes(M->kind[e]); ++a)
      for (int i = 0; i < 3; ++i)
        Ee += M->x[3 * M->T[4 * e + a] + i] * (-M->elem_mass[e] * g[i]);
    E += Ee;
    Eload += Ee;
  }
  /* EnergyElement_Force.cpp:43-47: E = x . f */
  if (M->has_f) {
    double Ef = 0;
    for (int k = 0; k < 3 * M->n; ++k) Ef += M->x[k] * cur_force(M, k);
    E += Ef;
    Eload += Ef;
  }
  M->escale = fabs(Eel) + fabs(Eload);
  return E;
}

void orc_model_gradient(orc_model* M, double* gout) {
#pragma omp parallel for schedule(static)
  for (int e = 0; e < M->m; ++e) {
    double xe[12];
    gather_x(M, e, xe);
    elem_egh(M, e, xe, NULL, M->eg + 12 * e, NULL);
  }
  const int N = 3 * M->n;
  memset(gout, 0, sizeof(double) * N);
  for (int e = 0; e < M->m; ++e)
    for (int a = 0; a < 4; ++a)
      for (int i = 0; i < 3; ++i) gout[3 * M->T[4 * e + a] + i] += M->eg[12 * e + 3 * a + i];
  double g[3];
  cur_gravity(M, g);
  for (int e = 0; e < M->m; ++e)
    for (int a = 0; a < mass_nodes(M->kind[e]); ++a)
      for (int i = 0; i < 3; ++i) gout[3 * M->T[4 * e + a] + i] += -M->elem_mass[e] * g[i];
  if (M->has_f)
    for (int k = 0; k < N; ++k) gout[k] += cur_force(M, k);
  /* Simulable.cpp:735-744 FixVectorDoFs */
  for (int k = 0; k < N; ++k)
    if (M->fixed[k]) gout[k] = 0;
}

long orc_model_nnz_lower(orc_model* M) { return M->nnz; }

double orc_model_energy_scale(const orc_model* M) { return M->escale; }

void orc_model_hessian_lower(orc_model* M, long* rowptr, int* col,
                             double* val) {
#pragma omp parallel for schedule(static)
  for (int e = 0; e < M->m; ++e) {
    double xe[12];
    gather_x(M, e, xe);
    elem_egh(M, e, xe, NULL, NULL, M->eH + 144 * e);
  }
  const int N = 3 * M->n;
  memcpy(rowptr, M->rowptr, sizeof(long) * (N + 1));
  memcpy(col, M->col, sizeof(int) * M->nnz);
  memset(val, 0, sizeof(double) * M->nnz);
  for (int e = 0; e < M->m; ++e) {
    const long* s = M->slot + 144L * e;
    const double* He = M->eH + 144L * e;
    for (int q = 0; q < 144; ++q)
      if (s[q] >= 0) val[s[q]] += He[q];
  }
  /* Simulable.cpp:746-760 FixMatrixDoFs */
  for (int r = 0; r < N; ++r)
    for (long p = rowptr[r]; p < rowptr[r + 1]; ++p) {
      const int c = col[p];
      if (M->fixed[r] || M->fixed[c]) val[p] = (r == c) ? 1.0 : 0.0;
    }
}

void orc_model_element_egh(orc_model* M, double* E, double* g, double* H) {
#pragma omp parallel for schedule(static)
  for (int e = 0; e < M->m; ++e) {
    double xe[12];
    gather_x(M, e, xe);
    elem_egh(M, e, xe, E ? E + e : NULL, g ? g + 12L * e : NULL, H ? H + 144L * e : NULL);
  }
}

/* ======================================================================= */
/* Conjugate gradient (Eigen ConjugateGradient, UpLo = Lower, Jacobi        */
/* preconditioner) -- textbook PCG (Saad, Iterative Methods, Alg. 9.1) with */
/* Eigen's stopping rule |r|^2 < max(tol^2 |b|^2, DBL_MIN) and warm start.   */
/* ======================================================================= */

/* y = A x using only stored entries with row >= col (selfadjoint lower) */
static void symv_lower(int n, const long* rp, const int* ci, const double* v,
                       const double* x, double* y) {
  memset(y, 0, sizeof(double) * n);
  for (int r = 0; r < n; ++r) {
    double acc = 0;
    for (long p = rp[r]; p < rp[r + 1]; ++p) {
      const int c = ci[p];
      if (c < r) {
        acc += v[p] * x[c];
        y[c] += v[p] * x[r];
      } else if (c == r) {
        acc += v[p] * x[r];
      }
    }
    y[r] += acc;
  }
}

static double dot(int n, const double* a, const double* b) {
  double s = 0;
  for (int i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

int orc_cg(int n, const long* rp, const int* ci, const double* v,
           const double* b, double* x, double tol, int maxit, double* err) {
  double* r = (double*)malloc(sizeof(double) * n);
  double* z = (double*)malloc(sizeof(double) * n);
  double* p = (double*)malloc(sizeof(double) * n);
  double* q = (double*)malloc(sizeof(double) * n);
  double* dinv = (double*)malloc(sizeof(double) * n);
  for (int i = 0; i < n; ++i) dinv[i] = 1.0;
  for (int r0 = 0; r0 < n; ++r0)
    for (long k = rp[r0]; k < rp[r0 + 1]; ++k)
      if (ci[k] == r0 && v[k] != 0.0) dinv[r0] = 1.0 / v[k];
  symv_lower(n, rp, ci, v, x, q);
  for (int i = 0; i < n; ++i) r[i] = b[i] - q[i];
  const double bb = dot(n, b, b);
  int it = 0;
  if (bb == 0.0) {
    for (int i = 0; i < n; ++i) x[i] = 0;
    *err = 0;
    goto done;
  }
  {
    double thr = tol * tol * bb;
    if (thr < 2.2250738585072014e-308) thr = 2.2250738585072014e-308;
    double rr = dot(n, r, r);
    if (rr < thr) {
      *err = sqrt(rr / bb);
      goto done;
    }
    for (int i = 0; i < n; ++i) p[i] = dinv[i] * r[i];
    double rz = dot(n, r, p);
    while (it < maxit) {
      symv_lower(n, rp, ci, v, p, q);
      const double alpha = rz / dot(n, p, q);
      for (int i = 0; i < n; ++i) {
        x[i] += alpha * p[i];
        r[i] -= alpha * q[i];
      }
      rr = dot(n, r, r);
      if (rr < thr) break;
      for (int i = 0; i < n; ++i) z[i] = dinv[i] * r[i];
      const double rz_old = rz;
      rz = dot(n, r, z);
      const double beta = rz / rz_old;
      for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
      ++it;
    }
    *err = sqrt(rr / bb);
  }
done:
  free(r);
  free(z);
  free(p);
  free(q);
  free(dinv);
  return it;
}

int orc_linear_solve(int n, const long* rp, const int* ci, double* v,
                     const double* b, double* x, double tol, int maxit,
                     int reg_iters, int* tot_iters, int* reg_used) {
  /* LinearSolver.cpp:57-66: normA = Frobenius norm of the stored entries */
  double nA = 0;
  for (long k = 0; k < rp[n]; ++k) nA += v[k] * v[k];
  nA = sqrt(nA);
  const double regT = 1e-6 * nA;
  double* xprev = (double*)malloc(sizeof(double) * n);
  memcpy(xprev, x, sizeof(double) * n);
  int status = 1;
  *reg_used = 0;
  for (int i = 0; i < reg_iters; ++i) {
    if (i != 0) {
      /* LinearSolver_EigenCG.cpp:73 mA += R * regT * 10^(i-1), cumulative */
      const double s = regT * pow(10.0, i - 1);
      for (int r0 = 0; r0 < n; ++r0)
        for (long k = rp[r0]; k < rp[r0 + 1]; ++k)
          if (ci[k] == r0) v[k] += s;
      *reg_used = i;
    }
    memcpy(x, xprev, sizeof(double) * n);
    double err;
    const int its = orc_cg(n, rp, ci, v, b, x, tol, maxit, &err);
    *tot_iters += its;
    if (dot(n, b, x) < 0.0) continue; /* :97-101 indefinite */
    status = 0;
    break;
  }
  free(xprev);
  return status;
}

int orc_sparse_ldlt_solve(int n, const long* rp, const int* ci, const double* v,
                          const double* b, double* x, double maxError, int regIters,
                          int* reg_used) {
  double nA = 0;
  for (long k = 0; k < rp[n]; ++k) nA += v[k] * v[k];
  double* A = (double*)calloc((size_t)n * n, sizeof(double));
  for (int r0 = 0; r0 < n; ++r0)
    for (long k = rp[r0]; k < rp[r0 + 1]; ++k)
      if (ci[k] <= r0) A[(long)r0 * n + ci[k]] = v[k];
  const int st = orc_ldlt_solve_norm(n, A, sqrt(nA), b, x, maxError, regIters, reg_used);
  free(A);
  return st;
}

/* ======================================================================= */
/* Newton (OptimSolver + OptimSolver_USQP_LS + OptimProblem_BasicStatic)    */
/* ======================================================================= */

void orc_newton_default_opts(orc_newton_opts* o) {
  /* OptimSolver.h:47-70 */
  o->tolMaxError = 1e-6;
  o->tolMinStep = 1e-5;
  o->minStepMaxCount = 3;
  o->tolMinImpr = 1e-9;
  o->minImprMaxCount = 3;
  o->numMaxIters = 100;
  o->trySteepest = 1;
  o->maxStepSize = 1000;
  o->lsNumRegTrials = 10;
  o->lsNumMaxIters = 1000;
  o->lsTolMaxError = 1e-9;
  o->lsearch_numTry = 32;
  o->lSearch_alpha = 0.5;
  o->ls_noise_guard = 0.0;
  o->ls_solver = 1;
  o->qpSolverType = 1; /* QPSolverType::Newton */
}

void orc_lbfgs_direction(int N, int M, int count, int newest, const double* S,
                         const double* Y, const double* g, double* dx) {
  double a[64];
  for (int k = 0; k < N; ++k) dx[k] = -g[k];
  if (count <= 0) return;
  /* newest to oldest */
  int j = newest;
  for (int i = 0; i < count; ++i) {
    const double* s = S + (size_t)j * N;
    const double* y = Y + (size_t)j * N;
    a[j] = dot(N, s, dx) / dot(N, y, s);
    for (int k = 0; k < N; ++k) dx[k] -= a[j] * y[k];
    if (i + 1 < count) j = (j + M - 1) % M;
  }
  {
    const double* s = S + (size_t)newest * N;
    const double* y = Y + (size_t)newest * N;
    const double gam = dot(N, s, y) / dot(N, y, y);
    for (int k = 0; k < N; ++k) dx[k] *= gam;
  }
  /* oldest to newest */
  for (int i = 0; i < count; ++i) {
    const double* s = S + (size_t)j * N;
    const double* y = Y + (size_t)j * N;
    const double beta = dot(N, y, dx) / dot(N, y, s);
    for (int k = 0; k < N; ++k) dx[k] += (a[j] - beta) * s[k];
    j = (j + 1) % M;
  }
}

static double norm2(int n, const double* a) { return sqrt(dot(n, a, a)); }

/* OptimSolver_USQP_LS::SolveStepInternal (OptimSolver_USQP_LS.cpp:18-68),
 * ComputeStep (OptimSolver.cpp:313-352), LineSearch_Simple (:688-736) and
 * UpdateStepResult (:402-466). */
#define ORC_LBFGS_M 3 /* the reference's history length (OptimSolver.cpp:527) */
typedef struct {
  int N;
  double *vx, *v0, *vg, *dx, *dxo, *vgNew, *xsave, *b, *xt;
  double obj, opt;
  int result, minImprCount, minStepCount;
  int lin_its, bis;
  /* LBFGS: previous x and g of the last LBFGS step, ring of pairs */
  int bfgs_init, bfgs_count, bfgs_newest;
  double *bx_p, *bg_p, *bS, *bY;
} nstate;

static void solve_step(orc_problem* P, const orc_newton_opts* o, nstate* S,
                       int steepest) {
  const int N = S->N;
  P->step_load(P->p);                 /* OnIterStart: StepBoundaryConditions */
  P->get_x(P->p, S->xsave);           /*              GetState */
  S->obj = P->energy(P->p);
  P->gradient(P->p, S->vg);
  memcpy(S->vx, S->xsave, sizeof(double) * N);
  S->opt = norm2(N, S->vg);
  if (S->opt < o->tolMaxError) {
    S->result = ORC_OS_SUCCESS;
    return;
  }
  if (!steepest && o->qpSolverType == 2) {
    /* ComputeStep_LBFGS (OptimSolver.cpp:522-591), corrected */
    if (!S->bfgs_init) {
      memcpy(S->bx_p, S->vx, sizeof(double) * N);
      memcpy(S->bg_p, S->vg, sizeof(double) * N);
      for (int k = 0; k < N; ++k) S->dx[k] = -S->vg[k];
      S->bfgs_init = 1;
    } else {
      const int slot = S->bfgs_count == 0 ? 0 : (S->bfgs_newest + 1) % ORC_LBFGS_M;
      double* sv = S->bS + (size_t)slot * N;
      double* yv = S->bY + (size_t)slot * N;
      for (int k = 0; k < N; ++k) {
        sv[k] = S->vx[k] - S->bx_p[k];
        yv[k] = S->vg[k] - S->bg_p[k];
      }
      memcpy(S->bx_p, S->vx, sizeof(double) * N);
      memcpy(S->bg_p, S->vg, sizeof(double) * N);
      S->bfgs_newest = slot;
      if (S->bfgs_count < ORC_LBFGS_M) S->bfgs_count++;
      orc_lbfgs_direction(N, ORC_LBFGS_M, S->bfgs_count, slot, S->bS, S->bY, S->vg, S->dx);
    }
  } else if (!steepest && o->qpSolverType != 0) {
    for (int k = 0; k < N; ++k) S->dx[k] = S->vx[k] - S->v0[k]; /* :488 */
    for (int k = 0; k < N; ++k) S->b[k] = -S->vg[k];
    S->lin_its += P->solve(P->p, S->b, S->dx, o);
  } else {
    for (int k = 0; k < N; ++k) S->dx[k] = -S->vg[k];
  }
  const double len = norm2(N, S->dx);
  if (len > o->maxStepSize)
    for (int k = 0; k < N; ++k) S->dx[k] = S->dx[k] / len * o->maxStepSize;
  double factor = 1.0, objNew = 0;
  int numBis = -1, guarded = 0;
  do {
    numBis++;
    for (int k = 0; k < N; ++k) S->dxo[k] = factor * S->dx[k];
    for (int k = 0; k < N; ++k) S->xt[k] = S->xsave[k] + S->dxo[k];
    P->set_x(P->p, S->xt);
    objNew = P->energy(P->p);
    P->gradient(P->p, S->vgNew);
    if (objNew < S->obj) break;
    if (o->ls_noise_guard > 0 &&
        objNew <= S->obj + o->ls_noise_guard * P->energy_scale(P->p) &&
        norm2(N, S->vgNew) < S->opt) {
      /* energy change below rounding: accept on the gradient norm and book
       * it as an improvement of zero (the decrease branch below) */
      guarded = 1;
      break;
    }
    factor *= o->lSearch_alpha;
    P->set_x(P->p, S->xsave); /* PrePerformStep */
  } while (norm2(N, S->dxo) > o->tolMinStep && numBis <= o->lsearch_numTry);
  S->bis += numBis;
  double impr = S->obj - objNew;
  const double step = norm2(N, S->dxo);
  if (guarded && impr < 0) impr = 0;
  if (impr > 0 || guarded) {
    const double optNew = norm2(N, S->vgNew);
    if (impr < o->tolMinImpr) {
      S->minImprCount++;
      S->result = ORC_OS_MINIMPR;
    } else if (step < o->tolMinStep) {
      S->minStepCount++;
      S->result = ORC_OS_MINSTEP;
    } else if (optNew <= o->tolMaxError) {
      S->result = ORC_OS_SUCCESS;
      S->minImprCount = 0;
      S->minStepCount = 0;
    } else {
      S->result = ORC_OS_ONGOING;
      S->minImprCount = 0;
      S->minStepCount = 0;
    }
    memcpy(S->v0, S->vx, sizeof(double) * N);
    for (int k = 0; k < N; ++k) S->vx[k] += S->dxo[k];
    memcpy(S->vg, S->vgNew, sizeof(double) * N);
    S->obj = objNew;
    S->opt = optNew;
  } else {
    S->result = ORC_OS_NONDESC;
  }
  P->iter_finish(P->p); /* OnIterFinish: UpdateKinematics */
}

/* OptimSolver::Init + SolveFull (OptimSolver.cpp:36-55,163-264) */
void orc_newton_generic(orc_problem* P, const orc_newton_opts* o,
                        orc_newton_result* R) {
  nstate S;
  memset(&S, 0, sizeof(S));
  const int N = P->N;
  S.N = N;
  double** bufs[] = {&S.vx, &S.v0, &S.vg, &S.dx, &S.dxo, &S.vgNew, &S.xsave, &S.b, &S.xt};
  for (int i = 0; i < 9; ++i) *bufs[i] = (double*)calloc(N, sizeof(double));
  S.bx_p = (double*)calloc(N, sizeof(double));
  S.bg_p = (double*)calloc(N, sizeof(double));
  S.bS = (double*)calloc((size_t)ORC_LBFGS_M * N, sizeof(double));
  S.bY = (double*)calloc((size_t)ORC_LBFGS_M * N, sizeof(double));
  P->get_x(P->p, S.vx);
  memcpy(S.v0, P->v0_keep ? P->v0_keep : S.vx, sizeof(double) * N);
  S.obj = P->energy(P->p);
  P->gradient(P->p, S.vg);
  S.opt = norm2(N, S.vg);
  S.result = ORC_OS_SUCCESS;
  int steepest = 0, itCount = 0;
  while (1) {
    itCount++;
    solve_step(P, o, &S, steepest);
    P->gradient(P->p, S.vg);
    S.opt = norm2(N, S.vg);
    if (P->fully_loaded(P->p)) {
      if (S.result == ORC_OS_NONDESC) {
        if (!o->trySteepest) break;
        if (steepest) break;
        steepest = 1;
      }
      if (S.result == ORC_OS_MINIMPR) {
        if (S.minImprCount == o->minImprMaxCount) break;
        if (o->trySteepest) steepest = !steepest;
      }
      if (S.result == ORC_OS_MINSTEP) {
        if (S.minStepCount == o->minStepMaxCount) break;
        if (o->trySteepest) steepest = !steepest;
      }
      if (S.result == ORC_OS_SUCCESS) {
        if (S.opt <= o->tolMaxError) break;
        S.result = ORC_OS_ONGOING;
      }
      if (S.result == ORC_OS_ONGOING) steepest = 0;
    }
    if (itCount + 1 > o->numMaxIters) {
      S.result = ORC_OS_MAXITER;
      break;
    }
  }
  R->result = S.result;
  R->iterations = itCount;
  R->cg_iterations = S.lin_its;
  R->line_search_bisections = S.bis;
  R->objective = S.obj;
  R->optimality = S.opt;
  if (P->v0_keep) memcpy(P->v0_keep, S.v0, sizeof(double) * N);
  for (int i = 0; i < 9; ++i) free(*bufs[i]);
  free(S.bx_p);
  free(S.bg_p);
  free(S.bS);
  free(S.bY);
}

/* tet model as a problem: Hessian in lower CSR + Jacobi CG with the
 * reference's regularization loop */
typedef struct {
  orc_model* M;
  long* rp;
  int* ci;
  double* val;
} tet_prob;
static double tp_energy(void* p) { return orc_model_energy(((tet_prob*)p)->M); }
static double tp_escale(void* p) { return orc_model_energy_scale(((tet_prob*)p)->M); }
static void tp_gradient(void* p, double* g) { orc_model_gradient(((tet_prob*)p)->M, g); }
static int tp_solve(void* p, const double* b, double* dx, const orc_newton_opts* o) {
  tet_prob* T = (tet_prob*)p;
  orc_model_hessian_lower(T->M, T->rp, T->ci, T->val);
  int its = 0, reg = 0;
  if (o->ls_solver == 2) {
    orc_sparse_ldlt_solve(3 * T->M->n, T->rp, T->ci, T->val, b, dx, o->lsTolMaxError,
                          o->lsNumRegTrials, &reg);
    return 0;
  }
  orc_linear_solve(3 * T->M->n, T->rp, T->ci, T->val, b, dx, o->lsTolMaxError,
                   o->lsNumMaxIters, o->lsNumRegTrials, &its, &reg);
  return its;
}
static int tp_step(void* p) { return orc_model_step_load(((tet_prob*)p)->M); }
static int tp_loaded(void* p) { return orc_model_fully_loaded(((tet_prob*)p)->M); }
static void tp_get(void* p, double* x) { orc_model_get_x(((tet_prob*)p)->M, x); }
static void tp_set(void* p, const double* x) { orc_model_set_x(((tet_prob*)p)->M, x); }
static void tp_finish(void* p) { (void)p; }

void orc_newton_solve(orc_model* M, const orc_newton_opts* o, orc_newton_result* R) {
  tet_prob T;
  T.M = M;
  T.rp = (long*)malloc(sizeof(long) * (3 * M->n + 1));
  T.ci = (int*)malloc(sizeof(int) * M->nnz);
  T.val = (double*)malloc(sizeof(double) * M->nnz);
  orc_problem P = {&T,      3 * M->n,  tp_energy, tp_escale, tp_gradient, tp_solve,
                   tp_step, tp_loaded, tp_get,    tp_set,    tp_finish, NULL};
  orc_newton_generic(&P, o, R);
  free(T.rp);
  free(T.ci);
  free(T.val);
}

/* ======================================================================= */
/* Implicit dynamics (OptimProblem_BasicDynamic.cpp:20-170)                 */
/* ======================================================================= */
struct orc_dynamic {
  orc_model* M;
  int N;
  double dt;
  double *v, *x0, *v0, *mass, *v0_keep;
  int v0_valid;
  long* rp;
  int* ci;
  double* val;
};

orc_dynamic* orc_dynamic_create(orc_model* M) {
  orc_dynamic* D = (orc_dynamic*)calloc(1, sizeof(orc_dynamic));
  D->M = M;
  D->N = 3 * M->n;
  D->v = (double*)calloc(D->N, sizeof(double));
  D->x0 = (double*)calloc(D->N, sizeof(double));
  D->v0 = (double*)calloc(D->N, sizeof(double));
  D->mass = (double*)calloc(D->N, sizeof(double));
  D->v0_keep = (double*)calloc(D->N, sizeof(double));
  D->rp = (long*)malloc(sizeof(long) * (D->N + 1));
  D->ci = (int*)malloc(sizeof(int) * M->nnz);
  D->val = (double*)malloc(sizeof(double) * M->nnz);
  orc_model_get_mass(M, D->mass); /* GetInertia: lumped, per DoF */
  return D;
}

void orc_dynamic_destroy(orc_dynamic* D) {
  if (!D) return;
  free(D->v);
  free(D->x0);
  free(D->v0);
  free(D->mass);
  free(D->v0_keep);
  free(D->rp);
  free(D->ci);
  free(D->val);
  free(D);
}

void orc_dynamic_set_velocity(orc_dynamic* D, const double* v) {
  memcpy(D->v, v, sizeof(double) * D->N);
}

void orc_dynamic_get_velocity(const orc_dynamic* D, double* v) {
  memcpy(v, D->v, sizeof(double) * D->N);
}

void orc_dynamic_begin(orc_dynamic* D, double dt) {
  D->dt = dt;
  if (!D->v0_valid) { /* OptimSolver::Init: m_v0 = x */
    memcpy(D->v0_keep, D->M->x, sizeof(double) * D->N);
    D->v0_valid = 1;
  }
  memcpy(D->x0, D->M->x, sizeof(double) * D->N);
  memcpy(D->v0, D->v, sizeof(double) * D->N);
  orc_model_step_load(D->M);
  for (int k = 0; k < D->N; ++k) D->M->x[k] = D->x0[k] + dt * D->v0[k];
}

/* va = (x - x0) / dt^2 - v0 / dt */
static double dyn_va(const orc_dynamic* D, int k) {
  const double dt = D->dt;
  return (D->M->x[k] - D->x0[k]) / (dt * dt) - D->v0[k] / dt;
}

double orc_dynamic_energy(orc_dynamic* D) {
  double e = orc_model_energy(D->M);
  double q = 0.0;
  for (int k = 0; k < D->N; ++k) {
    const double a = dyn_va(D, k);
    q += a * D->mass[k] * a;
  }
  return e + D->dt * D->dt / 2.0 * q;
}

void orc_dynamic_gradient(orc_dynamic* D, double* g) {
  orc_model_gradient(D->M, g);
  for (int k = 0; k < D->N; ++k) g[k] += D->mass[k] * dyn_va(D, k); /* + DMvDt = 0 */
  for (int k = 0; k < D->N; ++k)
    if (D->M->fixed[k]) g[k] = 0.0; /* FixVectorDoFs */
}

void orc_dynamic_hessian_lower(orc_dynamic* D, long* rp, int* ci, double* val) {
  orc_model_hessian_lower(D->M, rp, ci, val);
  const double idt2 = 1.0 / (D->dt * D->dt);
  for (int r = 0; r < D->N; ++r)
    for (long p = rp[r]; p < rp[r + 1]; ++p)
      if (ci[p] == r) val[p] += D->mass[r] * idt2;
  for (int r = 0; r < D->N; ++r) /* FixMatrixDoFs */
    for (long p = rp[r]; p < rp[r + 1]; ++p)
      if (D->M->fixed[r] || D->M->fixed[ci[p]]) val[p] = (r == ci[p]) ? 1.0 : 0.0;
}

static double dp_energy(void* p) { return orc_dynamic_energy((orc_dynamic*)p); }
static double dp_escale(void* p) { return orc_model_energy_scale(((orc_dynamic*)p)->M); }
static void dp_gradient(void* p, double* g) { orc_dynamic_gradient((orc_dynamic*)p, g); }
static int dp_solve(void* p, const double* b, double* dx, const orc_newton_opts* o) {
  orc_dynamic* D = (orc_dynamic*)p;
  orc_dynamic_hessian_lower(D, D->rp, D->ci, D->val);
  int its = 0, reg = 0;
  if (o->ls_solver == 2) {
    orc_sparse_ldlt_solve(D->N, D->rp, D->ci, D->val, b, dx, o->lsTolMaxError,
                          o->lsNumRegTrials, &reg);
    return 0;
  }
  orc_linear_solve(D->N, D->rp, D->ci, D->val, b, dx, o->lsTolMaxError, o->lsNumMaxIters,
                   o->lsNumRegTrials, &its, &reg);
  return its;
}
static int dp_step(void* p) { (void)p; return 0; }   /* OnIterStart: no load step */
static int dp_loaded(void* p) { (void)p; return 1; } /* IsFullyConstrained */
static void dp_get(void* p, double* x) { orc_model_get_x(((orc_dynamic*)p)->M, x); }
static void dp_set(void* p, const double* x) { orc_model_set_x(((orc_dynamic*)p)->M, x); }
static void dp_finish(void* p) { (void)p; }

void orc_dynamic_solve(orc_dynamic* D, const orc_newton_opts* o, orc_newton_result* R) {
  orc_problem P = {D,       D->N,      dp_energy, dp_escale, dp_gradient, dp_solve,
                   dp_step, dp_loaded, dp_get,    dp_set,    dp_finish,   D->v0_keep};
  orc_newton_generic(&P, o, R);
}

void orc_dynamic_end(orc_dynamic* D) {
  for (int k = 0; k < D->N; ++k) D->v[k] = (D->M->x[k] - D->x0[k]) / D->dt;
}

void orc_dynamic_step(orc_dynamic* D, double dt, const orc_newton_opts* o,
                      orc_newton_result* R) {
  orc_dynamic_begin(D, dt);
  orc_dynamic_solve(D, o, R);
  orc_dynamic_end(D);
}
