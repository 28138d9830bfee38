/*
 * Host C code of libbfamx (operators, brick and vmaps, face connectivity,
 * VTK writer, octree forest and mortars) driven over a range of sizes, built
 * with gcc -fsanitize=address,undefined by tests/test_host_sanitize.py.  Any
 * out-of-bounds access, leak or undefined behaviour aborts the run.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "bfamx.h"
int main(int argc, char **argv)
{
  const char *vtk_path = argc > 1 ? argv[1] : "host_sanitize.vtu";
  /* operators */
  for (int N = 1; N <= 8; ++N)
  {
    double r[9], w[9];
    if (bfamx_lgl(N, r, w)) return 1;
  }
  for (int a = 1; a <= 8; ++a)
    for (int b = 1; b <= 8; ++b)
    {
      double *prj = malloc(sizeof(double) * 5 * 81), *mp = malloc(sizeof(double) * 5 * 81),
             *wmp = malloc(sizeof(double) * 5 * 81);
      int id;
      if (bfamx_mortar_ops(a, b, prj, mp, wmp, &id)) { printf("ops %d %d\n", a, b); return 1; }
      free(prj); free(mp); free(wmp);
    }
  /* brick + vmaps + face connectivity */
  for (int dim = 2; dim <= 3; ++dim)
    for (int N = 1; N <= 7; N += 3)
    {
      int n[3] = {3, 2, 4}, per[3] = {1, 0, 1};
      double L[3] = {1, 2, 3};
      int32_t K = n[0] * n[1] * (dim == 3 ? n[2] : 1);
      int nf = 2 * dim;
      int32_t *EToE = malloc(sizeof(int32_t) * K * nf), *lex = malloc(sizeof(int32_t) * K);
      int8_t *EToF = malloc(K * nf);
      double *x0 = malloc(sizeof(double) * K * dim), *h = malloc(sizeof(double) * K * dim);
      if (bfamx_brick(dim, n, per, L, 1, EToE, EToF, lex, x0, h)) { printf("brick\n"); return 1; }
      int n1 = N + 1, Nfp = dim == 3 ? n1 * n1 : n1;
      int32_t *vM = malloc(sizeof(int32_t) * K * nf * Nfp), *vP = malloc(sizeof(int32_t) * K * nf * Nfp);
      if (bfamx_build_vmaps(dim, N, K, EToE, EToF, vM, vP)) { printf("vmaps\n"); return 1; }
      int32_t *ne = malloc(sizeof(int32_t) * K * nf);
      uint8_t *nc = malloc(K * nf);
      if (bfamx_face_connectivity(dim, N, K, vM, vP, ne, nc)) { printf("conn\n"); return 1; }
      /* vtk */
      double rr[9], ww[9];
      bfamx_lgl(N, rr, ww);
      int Np = dim == 3 ? n1 * n1 * n1 : n1 * n1;
      double *f0 = calloc((size_t)K * Np, sizeof(double));
      const char *names[1] = {"p"};
      const double *fields[1] = {f0};
      const int32_t ids[3] = {0, 2, 7};
      for (int bin = 0; bin < 3; ++bin)
        for (int nw = 0; nw <= 4; nw += 2)
          if (bfamx_vtk_write(vtk_path, dim, N, K, x0, h, rr, 1, names, fields, 0.5, bin, nw, ids))
          {
            printf("vtk\n");
            return 1;
          }
      {
        char idx[4096];
        snprintf(idx, sizeof idx, "%s.pvtu", vtk_path);
        const char *pieces[2] = {vtk_path, vtk_path};
        if (bfamx_vtk_write_index(idx, 2, pieces, 1, names, 2))
        {
          printf("vtk index\n");
          return 1;
        }
      }
      free(f0); free(ne); free(nc); free(vM); free(vP); free(EToE); free(lex); free(EToF); free(x0); free(h);
    }
  /* octree */
  for (int levels = 0; levels <= 3; ++levels)
  {
    bfamx_octree_desc_t d = {4, levels, {0.5, 0.5, 0.5}, 0.3, {5, 4, 3, 3}};
    bfamx_octree_t *o;
    if (bfamx_octree_build(&d, &o)) { printf("octree\n"); return 1; }
    int32_t K, nm, Kl[4];
    int ns;
    bfamx_octree_info(o, &K, &ns, &nm, Kl);
    for (int s = 0; s < ns; ++s)
    {
      int N; int32_t Ks;
      bfamx_octree_sub(o, s, &N, &Ks, NULL, NULL, NULL, NULL, NULL, NULL);
      int32_t *E = malloc(sizeof(int32_t) * Ks * 6), *gid = malloc(sizeof(int32_t) * Ks);
      int8_t *F = malloc(Ks * 6), *gl = malloc(Ks * 6);
      double *x0 = malloc(sizeof(double) * Ks * 3), *h = malloc(sizeof(double) * Ks * 3);
      bfamx_octree_sub(o, s, &N, &Ks, E, F, gl, x0, h, gid);
      free(E); free(gid); free(F); free(gl); free(x0); free(h);
    }
    bfamx_mortar_t *m = malloc(sizeof(bfamx_mortar_t) * (nm ? nm : 1));
    bfamx_octree_mortars(o, m);
    printf("octree levels %d: K %d subs %d mortars %d\n", levels, K, ns, nm);
    free(m);
    bfamx_octree_free(o);
  }
  printf("harness done\n");
  return 0;
}
