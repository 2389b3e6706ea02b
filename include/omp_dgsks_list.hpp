/*
 * omp_dgsks_list.hpp -- C++ batched interface, source-compatible with the
 * reference header (include/omp_dgsks_list.hpp:48-115 of ChenhanYu/ks): the
 * same five functions with the same parameter lists.
 *
 * In the reference these run an OpenMP task farm of 68 workers, one dgsks()
 * per task (frame/omp_dgsks_list.cpp:297-420).  Here each call flattens the
 * lists and runs all tasks in one persistent GPU launch (dgsks_list in ks.h).
 * Semantics kept: u[ulist[t][i]] accumulates task t's row i; alist, blist and
 * wlist must have equal sizes (else a message and exit(1), as in the
 * reference); empty lists return early; the *_unsymmetric / *_symmetric
 * variants compute the squared norms themselves (on the GPU), the symmetric
 * ones use XA as the source table and the unified variants use alist as ulist.
 */
#ifndef __OMP_DGSKS_LIST_HPP__
#define __OMP_DGSKS_LIST_HPP__

#include <stdio.h>
#include <stdlib.h>

#include <vector>

#include "ks.h"

void omp_dgsks_list_unsymmetric(
    ks_t   *kernel,
    int    k,
    std::vector<double> &u,
    int    nxa,
    double *XA,
    std::vector< std::vector<int> > &alist,
    int    nxb,
    double *XB,
    std::vector< std::vector<int> > &blist,
    double *w,
    std::vector< std::vector<int> > &wlist
    );

void omp_dgsks_list_symmetric(
    ks_t   *kernel,
    int    k,
    std::vector<double> &u,
    double *XA,
    int    nxa,
    std::vector< std::vector<int> > &alist,
    std::vector< std::vector<int> > &blist,
    double *w,
    std::vector< std::vector<int> > &wlist
    );

void omp_dgsks_list_separated_u_unsymmetric(
    ks_t   *kernel,
    int    k,
    std::vector<double> &u,
    std::vector< std::vector<int> > &ulist,
    int    nxa,
    double *XA,
    std::vector< std::vector<int> > &alist,
    int    nxb,
    double *XB,
    std::vector< std::vector<int> > &blist,
    double *w,
    std::vector< std::vector<int> > &wlist
    );

void omp_dgsks_list_separated_u_symmetric(
    ks_t   *kernel,
    int    k,
    std::vector<double> &u,
    std::vector< std::vector<int> > &ulist,
    double *XA,
    int    nxa,
    std::vector< std::vector<int> > &alist,
    std::vector< std::vector<int> > &blist,
    double *w,
    std::vector< std::vector<int> > &wlist
    );

void omp_dgsks_list(
    ks_t   *kernel,
    int    k,
    std::vector<double> &u,
    std::vector< std::vector<int> > &ulist,
    double *XA,
    double *XA2,
    std::vector< std::vector<int> > &alist,
    double *XB,
    double *XB2,
    std::vector< std::vector<int> > &blist,
    double *w,
    std::vector< std::vector<int> > &wlist
    );

#endif
