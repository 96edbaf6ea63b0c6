/* wm_mpi_stub.c -- the MPI routines the reference's apps and utilities call,
 * for one process (MPI_COMM_WORLD of size 1).  The Fortran callers use the
 * mpif.h convention: external procedures with every argument by reference,
 * trailing underscore.  Collectives over one rank copy the send buffer to
 * the receive buffer; point-to-point exchanges are with the rank itself.
 * The datatype codes are the ones module mpi (mpi.f90) publishes. */
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

static size_t type_size(int t) {
  switch (t) {
    case 1: return 4;   /* MPI_INTEGER, MPI_INTEGER4 */
    case 2: return 8;   /* MPI_INTEGER8 */
    case 3: return 8;   /* MPI_DOUBLE_PRECISION, MPI_REAL8 */
    case 4: return 4;   /* MPI_REAL, MPI_REAL4 */
    case 5: return 1;   /* MPI_CHARACTER */
    case 6: return 4;   /* MPI_LOGICAL */
    case 7: return 1;   /* MPI_BYTE */
    default:
      fprintf(stderr, "mpi stub: unknown datatype %d\n", t);
      exit(1);
  }
}

static void copy(const void *s, void *r, int n, int t) {
  if (s != r && n > 0) memmove(r, s, (size_t)n * type_size(t));
}

void mpi_init_(int *ierr) { *ierr = 0; }
void mpi_finalize_(int *ierr) { *ierr = 0; }
void mpi_abort_(int *comm, int *code, int *ierr) {
  (void)comm;
  *ierr = 0;
  exit(*code);
}
void mpi_comm_size_(int *comm, int *size, int *ierr) {
  (void)comm;
  *size = 1;
  *ierr = 0;
}
void mpi_comm_rank_(int *comm, int *rank, int *ierr) {
  (void)comm;
  *rank = 0;
  *ierr = 0;
}
void mpi_barrier_(int *comm, int *ierr) {
  (void)comm;
  *ierr = 0;
}
void mpi_bcast_(void *buf, int *n, int *t, int *root, int *comm, int *ierr) {
  (void)buf; (void)n; (void)t; (void)root; (void)comm;
  *ierr = 0;
}
void mpi_reduce_(void *s, void *r, int *n, int *t, int *op, int *root, int *comm, int *ierr) {
  (void)op; (void)root; (void)comm;
  copy(s, r, *n, *t);
  *ierr = 0;
}
void mpi_allreduce_(void *s, void *r, int *n, int *t, int *op, int *comm, int *ierr) {
  (void)op; (void)comm;
  copy(s, r, *n, *t);
  *ierr = 0;
}
void mpi_allgather_(void *s, int *ns, int *ts, void *r, int *nr, int *tr, int *comm, int *ierr) {
  (void)nr; (void)tr; (void)comm;
  copy(s, r, *ns, *ts);
  *ierr = 0;
}
void mpi_gather_(void *s, int *ns, int *ts, void *r, int *nr, int *tr, int *root, int *comm,
                 int *ierr) {
  (void)nr; (void)tr; (void)root; (void)comm;
  copy(s, r, *ns, *ts);
  *ierr = 0;
}
void mpi_sendrecv_(void *s, int *ns, int *ts, int *dest, int *stag, void *r, int *nr, int *tr,
                   int *src, int *rtag, int *comm, int *status, int *ierr) {
  (void)nr; (void)tr; (void)dest; (void)stag; (void)src; (void)rtag; (void)comm; (void)status;
  copy(s, r, *ns, *ts);
  *ierr = 0;
}
double mpi_wtime_(void) {
  struct timeval tv;
  gettimeofday(&tv, NULL);
  return (double)tv.tv_sec + 1e-6 * (double)tv.tv_usec;
}
