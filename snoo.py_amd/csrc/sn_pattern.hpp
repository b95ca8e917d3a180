// Layout of the pattern handle shared by the scatter and CG translation units.
#pragma once

struct sn_pattern {
    int n = 0;                 // number of DOFs (rows)
    int num_el = 0;            // elements over all domains
    long long nnz = 0;
    int *indptr = nullptr;     // [n+1]
    int *indices = nullptr;    // [nnz], sorted within each row
    int *diag = nullptr;       // [n] position of (i,i), -1 if absent
    int *nz_ptr = nullptr;     // [nnz+1] ranges into contrib
    int *contrib = nullptr;    // [36*num_el] local pair slots pr*num_el + e
    int *dof_ptr = nullptr;    // [n+1] ranges into dof_contrib
    int *dof_contrib = nullptr;// [6*num_el] local DOF slots j*num_el + e
};
