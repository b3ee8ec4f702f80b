/*
 * setup_mpi.h -- the communicator set-up a reference main program calls
 * (src/setup_mpi.h:1-32, used by src/main.c:25-40,92), for the single-GPU
 * host facade.
 *
 * One process drives one GPU and owns the whole grid, so the facade keeps
 * the rank/size globals main() fills with MPI_Comm_rank / MPI_Comm_size and
 * the two set-up calls, and nothing else: there is no Cartesian communicator
 * and no halo datatypes (the device ghost fill replaces the exchange,
 * src/boundary_exchange.c).  setup_comm_and_types() returns -1 unless the
 * configuration asks for one subdomain (proc_dims 1 1 1) and this is rank 0
 * of a world of one; free_mpi_types() does nothing.
 *
 * <mpi.h> is included as in the reference because main() calls MPI_Init and
 * friends itself; the facade library does not link MPI.
 */
#ifndef SETUP_MPI_H
#define SETUP_MPI_H

#include <mpi.h>

/* rank and size of this process in MPI_COMM_WORLD (written by main()) */
extern int rank, size;

/* single subdomain check (src/setup_mpi.c:28-59 builds the communicator) */
int setup_comm_and_types(int rank);

/* nothing to free (src/setup_mpi.c:255-273) */
void free_mpi_types();

#endif
