/*
 * VTK XML unstructured-grid (.vtu) output of a subdomain's nodal fields.
 *
 * Writes what the reference writes per subdomain piece
 * (bfam_subdomain_dgx_vtk_write_vtu_piece, bfam.c:6050-6572, and the
 * resampling of bfam_subdomain_dgx_vtk_interp, bfam.c:5872-5929): the nodal
 * values, optionally resampled to Np_write equispaced points per direction
 * ("poor man's high-order viz"), every element split into N_vtk^dim linear
 * cells (VTK_PIXEL / VTK_VOXEL, vertices in lexicographic order), Int32
 * connectivity and offsets, the cell arrays "time", "mpirank",
 * "subdomain_id" and "root_id", and one point array per field.  Files are
 * version 0.1 without a header_type, so VTK reads 32-bit header words:
 * binary output is base64 of [uint32 byte count][data]; compressed output
 * (binary = 2, the reference's compressor="vtkZLibDataCompressor" files
 * written through sc_vtk_write_compressed, bfam.c:5931-5938, 8278-8281)
 * splits each array into 32 KiB blocks deflated by zlib, after a header of
 * 32-bit words: block count, block size, size of the last block, then each
 * block's compressed size.  The header and the blocks are base64-encoded
 * separately, as VTK reads them.  Affine elements: node coordinates come
 * from each element's lower corner and edge lengths.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

#include "bfamx_internal.h"

static const char b64[] =
    "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";

/* base64 of n bytes, padded to a multiple of 4 characters */
static void put_base64(FILE *fp, const unsigned char *buf, size_t total)
{
  for (size_t i = 0; i < total; i += 3)
  {
    const unsigned a = buf[i];
    const unsigned b = i + 1 < total ? buf[i + 1] : 0;
    const unsigned c = i + 2 < total ? buf[i + 2] : 0;
    const unsigned v = (a << 16) | (b << 8) | c;
    fputc(b64[(v >> 18) & 63], fp);
    fputc(b64[(v >> 12) & 63], fp);
    fputc(i + 1 < total ? b64[(v >> 6) & 63] : '=', fp);
    fputc(i + 2 < total ? b64[v & 63] : '=', fp);
  }
}

enum
{
  ZBLOCK = 1 << 15 /* uncompressed bytes per compressed block */
};

/* base64 of [uint32 nbytes][data], as VTK reads "binary" inline arrays of
 * version-0.1 files; with compress, the zlib block layout above.  Returns
 * 0, or -1 if memory or zlib fails or an array does not fit 32-bit words. */
static int write_base64(FILE *fp, const void *data, size_t nbytes, int compress)
{
  if (nbytes > UINT32_MAX)
    return -1;
  if (!compress)
  {
    const uint32_t hdr = (uint32_t)nbytes;
    unsigned char *buf = malloc(4 + nbytes);
    if (!buf)
      return -1;
    memcpy(buf, &hdr, 4);
    memcpy(buf + 4, data, nbytes);
    put_base64(fp, buf, 4 + nbytes);
    fputc('\n', fp);
    free(buf);
    return 0;
  }
  const size_t nblk = (nbytes + ZBLOCK - 1) / ZBLOCK;
  const size_t last = nbytes - (nblk ? (nblk - 1) * (size_t)ZBLOCK : 0);
  uint32_t *hdr = malloc(sizeof(uint32_t) * (3 + nblk));
  const uLong bound = compressBound(ZBLOCK);
  unsigned char *z = malloc(nblk * (size_t)bound + 1);
  if (!hdr || !z)
  {
    free(hdr);
    free(z);
    return -1;
  }
  hdr[0] = (uint32_t)nblk;
  hdr[1] = ZBLOCK;
  hdr[2] = (uint32_t)last;
  size_t zt = 0;
  for (size_t k = 0; k < nblk; ++k)
  {
    const size_t n = k + 1 < nblk ? (size_t)ZBLOCK : last;
    uLongf zn = bound;
    if (compress2(z + zt, &zn, (const Bytef *)data + k * (size_t)ZBLOCK, (uLong)n,
                  Z_BEST_COMPRESSION) != Z_OK)
    {
      free(hdr);
      free(z);
      return -1;
    }
    hdr[3 + k] = (uint32_t)zn;
    zt += zn;
  }
  put_base64(fp, (const unsigned char *)hdr, sizeof(uint32_t) * (3 + nblk));
  put_base64(fp, z, zt);
  fputc('\n', fp);
  free(hdr);
  free(z);
  return 0;
}

/* binary: 0 ascii, 1 base64, 2 zlib-compressed base64.  Each returns 0, or
 * -1 if compression fails. */
static int write_f64(FILE *fp, const char *name, int ncomp, const double *v, size_t n,
                     int binary)
{
  int rc = 0;
  fprintf(fp,
          "        <DataArray type=\"Float64\" Name=\"%s\" NumberOfComponents=\"%d\" "
          "format=\"%s\">\n",
          name, ncomp, binary ? "binary" : "ascii");
  if (binary)
    rc = write_base64(fp, v, sizeof(double) * n * (size_t)ncomp, binary == 2);
  else
    for (size_t i = 0; i < n * (size_t)ncomp; ++i)
      fprintf(fp, "%.17g\n", v[i]);
  fprintf(fp, "        </DataArray>\n");
  return rc;
}

static int write_i32(FILE *fp, const char *name, const int32_t *v, size_t n, int binary)
{
  int rc = 0;
  fprintf(fp, "        <DataArray type=\"Int32\" Name=\"%s\" format=\"%s\">\n", name,
          binary ? "binary" : "ascii");
  if (binary)
    rc = write_base64(fp, v, sizeof(int32_t) * n, binary == 2);
  else
    for (size_t i = 0; i < n; ++i)
      fprintf(fp, "%d\n", (int)v[i]);
  fprintf(fp, "        </DataArray>\n");
  return rc;
}

static int write_u8(FILE *fp, const char *name, const uint8_t *v, size_t n, int binary)
{
  int rc = 0;
  fprintf(fp, "        <DataArray type=\"UInt8\" Name=\"%s\" format=\"%s\">\n", name,
          binary ? "binary" : "ascii");
  if (binary)
    rc = write_base64(fp, v, n, binary == 2);
  else
    for (size_t i = 0; i < n; ++i)
      fprintf(fp, "%u\n", (unsigned)v[i]);
  fprintf(fp, "        </DataArray>\n");
  return rc;
}

int bfamx_vtk_write(const char *path, int dim, int N, int32_t K, const double *elem_x0,
                    const double *elem_h, const double *r, int nfields,
                    const char *const *names, const double *const *fields, double time,
                    int binary, int Np_write, const int32_t *ids)
{
  if (!path || (dim != 2 && dim != 3) || N < 1 || K < 0 || !elem_x0 || !elem_h || !r ||
      nfields < 0 || (nfields && (!names || !fields)) || Np_write == 1 || Np_write < 0 ||
      binary < 0 || binary > 2)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_vtk_write: bad arguments");
  const int n1 = N + 1;
  const int nv = Np_write > 0 ? Np_write : n1; /* points per direction */
  const int Npw = dim == 2 ? nv * nv : nv * nv * nv;
  const int Nps = dim == 2 ? n1 * n1 : n1 * n1 * n1;
  /* 1-D interpolation from the LGL nodes to the output points */
  double *I = malloc(sizeof(double) * (size_t)nv * n1);
  double *ref = malloc(sizeof(double) * nv);
  for (int p = 0; p < nv; ++p)
  {
    ref[p] = Np_write > 0 ? -1.0 + 2.0 * p / (nv - 1) : r[p];
    bfamx_lagrange_eval(N, r, ref[p], I + (size_t)p * n1);
  }
  const size_t npts = (size_t)K * Npw;
  const int cpe = dim == 2 ? (nv - 1) * (nv - 1) : (nv - 1) * (nv - 1) * (nv - 1);
  const size_t ncells = (size_t)K * cpe;
  const int nvert = dim == 2 ? 4 : 8;
  /* connectivity and offsets are Int32, as the reference's bfam_locidx_t */
  if ((double)nvert * (double)ncells > (double)INT32_MAX)
  {
    free(I);
    free(ref);
    return bfamx_fail(BFAMX_EINVAL, "bfamx_vtk_write: %zu cells do not fit Int32 offsets",
                      ncells);
  }
  double *xyz = malloc(sizeof(double) * 3 * npts);
  int32_t *conn = malloc(sizeof(int32_t) * nvert * ncells);
  int32_t *offs = malloc(sizeof(int32_t) * ncells);
  uint8_t *type = malloc(ncells);
  double *tcell = malloc(sizeof(double) * (ncells ? ncells : 1));
  int32_t *icell = malloc(sizeof(int32_t) * (ncells ? ncells : 1));
  double *val = malloc(sizeof(double) * (npts ? npts : 1));
  FILE *fp = fopen(path, "w");
  int rc = BFAMX_OK;
  if (!I || !ref || !xyz || !conn || !offs || !type || !tcell || !icell || !val || !fp)
  {
    rc = bfamx_fail(BFAMX_EINVAL, "bfamx_vtk_write: cannot write %s", path);
    goto done;
  }
  for (int32_t e = 0; e < K; ++e)
    for (int p = 0; p < Npw; ++p)
    {
      const int idx[3] = {p % nv, (p / nv) % nv, p / (nv * nv)};
      double *X = xyz + 3 * ((size_t)e * Npw + p);
      X[2] = 0;
      for (int d = 0; d < dim; ++d)
        X[d] = elem_x0[(size_t)e * dim + d] +
               0.5 * (ref[idx[d]] + 1.0) * elem_h[(size_t)e * dim + d];
    }
  {
    /* pixels / voxels with their vertices in lexicographic order, as the
     * reference lists them (bfam.c:6160-6210) */
    size_t c = 0;
    for (int32_t e = 0; e < K; ++e)
    {
      const int32_t base = e * Npw;
      const int nk = dim == 3 ? nv - 1 : 1;
      for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nv - 1; ++j)
          for (int i = 0; i < nv - 1; ++i, ++c)
          {
            int32_t *v = conn + (size_t)nvert * c;
            const int32_t o = base + i + nv * (j + nv * k);
            for (int t = 0; t < nvert; ++t)
              v[t] = o + (t & 1) + nv * ((t >> 1) & 1) + nv * nv * (t >> 2);
            offs[c] = (int32_t)(nvert * (c + 1));
            type[c] = dim == 2 ? 8 : 11; /* VTK_PIXEL, VTK_VOXEL */
            tcell[c] = time;
          }
    }
  }
  fprintf(fp, "<?xml version=\"1.0\"?>\n");
  fprintf(fp, "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\"%s byte_order=\"LittleEndian\">\n",
          binary == 2 ? " compressor=\"vtkZLibDataCompressor\"" : "");
  int zerr = 0;
  fprintf(fp, "  <UnstructuredGrid>\n");
  fprintf(fp, "    <Piece NumberOfPoints=\"%zu\" NumberOfCells=\"%zu\">\n", npts, ncells);
  fprintf(fp, "      <Points>\n");
  zerr |= write_f64(fp, "Position", 3, xyz, npts, binary);
  fprintf(fp, "      </Points>\n");
  fprintf(fp, "      <Cells>\n");
  zerr |= write_i32(fp, "connectivity", conn, (size_t)nvert * ncells, binary);
  zerr |= write_i32(fp, "offsets", offs, ncells, binary);
  zerr |= write_u8(fp, "types", type, ncells, binary);
  fprintf(fp, "      </Cells>\n");
  fprintf(fp, "      <CellData Scalars=\"time,mpirank,subdomain_id,root_id\">\n");
  zerr |= write_f64(fp, "time", 1, tcell, ncells, binary);
  {
    static const char *idname[3] = {"mpirank", "subdomain_id", "root_id"};
    for (int a = 0; a < 3; ++a)
    {
      for (size_t c = 0; c < ncells; ++c)
        icell[c] = ids ? ids[a] : 0;
      zerr |= write_i32(fp, idname[a], icell, ncells, binary);
    }
  }
  fprintf(fp, "      </CellData>\n");
  fprintf(fp, "      <PointData>\n");
  {
    /* sum factorisation: one 1-D interpolation per direction */
    const int nk = dim == 3 ? n1 : 1, mk = dim == 3 ? nv : 1;
    double *t1 = malloc(sizeof(double) * (size_t)nv * n1 * n1);
    double *t2 = malloc(sizeof(double) * (size_t)nv * nv * n1);
    for (int f = 0; f < nfields; ++f)
    {
      const double *s = fields[f];
      for (int32_t e = 0; e < K; ++e)
      {
        const double *u = s + (size_t)e * Nps;
        for (int k = 0; k < nk; ++k)
          for (int j = 0; j < n1; ++j)
            for (int a = 0; a < nv; ++a)
            {
              double acc = 0;
              for (int i = 0; i < n1; ++i)
                acc += I[(size_t)a * n1 + i] * u[i + n1 * (j + n1 * k)];
              t1[a + (size_t)nv * (j + (size_t)n1 * k)] = acc;
            }
        for (int k = 0; k < nk; ++k)
          for (int b = 0; b < nv; ++b)
            for (int a = 0; a < nv; ++a)
            {
              double acc = 0;
              for (int j = 0; j < n1; ++j)
                acc += I[(size_t)b * n1 + j] * t1[a + (size_t)nv * (j + (size_t)n1 * k)];
              t2[a + (size_t)nv * (b + (size_t)nv * k)] = acc;
            }
        double *o = val + (size_t)e * Npw;
        for (int c3 = 0; c3 < mk; ++c3)
          for (int b = 0; b < nv; ++b)
            for (int a = 0; a < nv; ++a)
            {
              double acc = 0;
              if (dim == 3)
                for (int k = 0; k < n1; ++k)
                  acc += I[(size_t)c3 * n1 + k] * t2[a + (size_t)nv * (b + (size_t)nv * k)];
              else
                acc = t2[a + (size_t)nv * b];
              o[a + (size_t)nv * (b + (size_t)nv * c3)] = acc;
            }
      }
      zerr |= write_f64(fp, names[f], 1, val, npts, binary);
    }
    free(t1);
    free(t2);
  }
  fprintf(fp, "      </PointData>\n");
  fprintf(fp, "    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n");
  if (zerr)
    rc = bfamx_fail(BFAMX_EINVAL, "bfamx_vtk_write: compression failed for %s", path);
  else if (ferror(fp))
    rc = bfamx_fail(BFAMX_EINVAL, "bfamx_vtk_write: write error on %s", path);
done:
  if (fp && fclose(fp) && !rc)
    rc = bfamx_fail(BFAMX_EINVAL, "bfamx_vtk_write: cannot close %s", path);
  free(I);
  free(ref);
  free(xyz);
  free(conn);
  free(offs);
  free(type);
  free(tcell);
  free(icell);
  free(val);
  return rc;
}

int bfamx_vtk_write_index(const char *path, int npieces, const char *const *piece_paths,
                          int nfields, const char *const *names, int binary)
{
  if (!path || npieces < 0 || (npieces && !piece_paths) || nfields < 0 ||
      (nfields && !names) || binary < 0 || binary > 2)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_vtk_write_index: bad arguments");
  for (int i = 0; i < npieces; ++i)
    if (!piece_paths[i])
      return bfamx_fail(BFAMX_EINVAL, "bfamx_vtk_write_index: piece %d is NULL", i);
  for (int f = 0; f < nfields; ++f)
    if (!names[f])
      return bfamx_fail(BFAMX_EINVAL, "bfamx_vtk_write_index: name %d is NULL", f);
  FILE *fp = fopen(path, "w");
  if (!fp)
    return bfamx_fail(BFAMX_EINVAL, "bfamx_vtk_write_index: cannot write %s", path);
  const char *fmt = binary ? "binary" : "ascii";
  fprintf(fp, "<?xml version=\"1.0\"?>\n");
  fprintf(fp, "<VTKFile type=\"PUnstructuredGrid\" version=\"0.1\"%s byte_order=\"LittleEndian\">\n",
          binary == 2 ? " compressor=\"vtkZLibDataCompressor\"" : "");
  fprintf(fp, "  <PUnstructuredGrid GhostLevel=\"0\">\n");
  fprintf(fp, "    <PPoints>\n");
  fprintf(fp,
          "      <PDataArray type=\"Float64\" Name=\"Position\" NumberOfComponents=\"3\" "
          "format=\"%s\"/>\n",
          fmt);
  fprintf(fp, "    </PPoints>\n");
  fprintf(fp, "    <PCellData Scalars=\"time,mpirank,subdomain_id,root_id\">\n");
  fprintf(fp, "      <PDataArray type=\"Float64\" Name=\"time\" format=\"%s\"/>\n", fmt);
  fprintf(fp, "      <PDataArray type=\"Int32\" Name=\"mpirank\" format=\"%s\"/>\n", fmt);
  fprintf(fp, "      <PDataArray type=\"Int32\" Name=\"subdomain_id\" format=\"%s\"/>\n", fmt);
  fprintf(fp, "      <PDataArray type=\"Int32\" Name=\"root_id\" format=\"%s\"/>\n", fmt);
  fprintf(fp, "    </PCellData>\n");
  fprintf(fp, "    <PPointData>\n");
  for (int f = 0; f < nfields; ++f)
    fprintf(fp, "      <PDataArray type=\"Float64\" Name=\"%s\" format=\"%s\"/>\n",
            names[f], fmt);
  fprintf(fp, "    </PPointData>\n");
  for (int i = 0; i < npieces; ++i)
    fprintf(fp, "    <Piece Source=\"%s\"/>\n", piece_paths[i]);
  fprintf(fp, "  </PUnstructuredGrid>\n</VTKFile>\n");
  int rc = BFAMX_OK;
  if (ferror(fp))
    rc = bfamx_fail(BFAMX_EINVAL, "bfamx_vtk_write_index: write error on %s", path);
  if (fclose(fp) && !rc)
    rc = bfamx_fail(BFAMX_EINVAL, "bfamx_vtk_write_index: cannot close %s", path);
  return rc;
}
