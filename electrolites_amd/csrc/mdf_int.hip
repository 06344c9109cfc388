// C-ABI entry points of the Rys integral kernels.
//
// mdf_int3c2e fills one pair-chunk of the three-centre tensor in the layout
// the metric solve consumes: out[P][p - p0], P over spherical aux functions,
// p = i(i+1)/2 + j (i >= j) over the chunk's packed pairs.  It plays the part
// of GPU4PySCF's int3c2e evaluator in the reference's cderi build
// (electrolites/fastdfcderi.py:404-409, called at :604 and :667); the aux
// functions come out spherical and contracted, so no separate cart->sph /
// recontraction pass (electrolites/_dfaux.py:75-119) is needed.
// mdf_int2c2e fills the metric (P|Q) (electrolites/_dfj2c.py:234-262).
#include "mdf_int.h"
#include "../../include/mi355df.h"

using namespace mdf;

static BasisDev to_dev(const mdf_basis *b) {
    BasisDev d;
    d.nprim = b->nprim;
    d.poff = b->poff;
    d.aooff = b->aooff;
    d.xyz = b->xyz;
    d.exps = b->exps;
    d.coefs = b->coefs;
    return d;
}

extern "C" int mdf_int3c2e(const mdf_basis *ao, const mdf_basis *aux, const int *pairs,
                           const long *pair_class_off, const int *aux_sorted, const long *aux_l_off,
                           long p0, const int *col_map, double *out, long ldo, void *stream) {
    MDF_CHECK_ARG(ao && aux && pair_class_off && aux_l_off && out, "mdf_int3c2e: null argument");
    MDF_CHECK_ARG(ao->lmax <= LMAX_AO && aux->lmax <= LMAX_AUX,
                  "mdf_int3c2e: angular momentum beyond the generated classes (ao %d > %d or aux %d > %d)",
                  ao->lmax, LMAX_AO, aux->lmax, LMAX_AUX);
    hipStream_t st = as_stream(stream);
    const BasisDev a = to_dev(ao), x = to_dev(aux);
    for (int la = 0; la <= LMAX_AO; la++)
        for (int lb = 0; lb <= LMAX_AO; lb++) {
            const int cls = la * (LMAX_AO + 1) + lb;
            const long q0 = pair_class_off[cls], q1 = pair_class_off[cls + 1];
            if (q1 <= q0) continue;
            MDF_CHECK_ARG(q1 - q0 < (1L << 31), "mdf_int3c2e: too many pairs in one class");
            for (int lc = 0; lc <= LMAX_AUX; lc++) {
                const long k0 = aux_l_off[lc], k1 = aux_l_off[lc + 1];
                if (k1 <= k0) continue;
                int rc = launch_int3c(la, lb, lc, a, x, pairs + 2 * q0, (int)(q1 - q0), aux_sorted + k0,
                                      (int)(k1 - k0), out, ldo, p0, col_map, st);
                if (rc) return rc;
            }
        }
    return MDF_OK;
}

extern "C" int mdf_int2c2e(const mdf_basis *aux, const int *aux_sorted, const long *aux_l_off, double *out,
                           long ldo, void *stream) {
    MDF_CHECK_ARG(aux && aux_l_off && out, "mdf_int2c2e: null argument");
    MDF_CHECK_ARG(aux->lmax <= LMAX_AUX, "mdf_int2c2e: aux angular momentum %d > %d", aux->lmax, LMAX_AUX);
    hipStream_t st = as_stream(stream);
    const BasisDev x = to_dev(aux);
    for (int la = 0; la <= LMAX_AUX; la++) {
        const long p0 = aux_l_off[la], p1 = aux_l_off[la + 1];
        if (p1 <= p0) continue;
        for (int lc = 0; lc <= LMAX_AUX; lc++) {
            const long k0 = aux_l_off[lc], k1 = aux_l_off[lc + 1];
            if (k1 <= k0) continue;
            int rc = launch_int2c(la, lc, x, aux_sorted + p0, (int)(p1 - p0), aux_sorted + k0, (int)(k1 - k0),
                                  out, ldo, st);
            if (rc) return rc;
        }
    }
    return MDF_OK;
}

// Nuclear derivative of the three-centre term of the DF gradient:
// grad[3 * atom + x] += sum_{P, i, j} W[P, (i, j)] d(ij|P)/dR_atom,x over the
// orbital centres of the listed shell pairs (both orders of an off-diagonal
// pair), and the aux centre's share (translational invariance) into
// grad_aux.  Plays the part of GPU4PySCF's sum_ejk_int3c2e_ip1 in the
// reference's gradient (electrolites/fastdfgrad.py:1014-1040).
extern "C" int mdf_int3c2e_ip(const mdf_basis *ao, const mdf_basis *aux, const int *pairs,
                              const long *pair_class_off, const int *aux_sorted, const long *aux_l_off,
                              const double *W, long ldw, const int *col_map, const int *ao_atom,
                              const int *aux_atom, int natm, double *grad, double *grad_aux, void *stream) {
    MDF_CHECK_ARG(ao && aux && pair_class_off && aux_l_off && W && ao_atom && aux_atom && grad && grad_aux,
                  "mdf_int3c2e_ip: null argument");
    MDF_CHECK_ARG(natm > 0, "mdf_int3c2e_ip: natm must be positive");
    MDF_CHECK_ARG(ao->lmax <= LMAX_AO && aux->lmax <= LMAX_AUX,
                  "mdf_int3c2e_ip: angular momentum beyond the generated classes (ao %d > %d or aux %d > %d)",
                  ao->lmax, LMAX_AO, aux->lmax, LMAX_AUX);
    MDF_CHECK_ARG(ldw > 0, "mdf_int3c2e_ip: ldw must be positive");
    hipStream_t st = as_stream(stream);
    const BasisDev a = to_dev(ao), x = to_dev(aux);
    for (int la = 0; la <= LMAX_AO; la++)
        for (int lb = 0; lb <= LMAX_AO; lb++) {
            const int cls = la * (LMAX_AO + 1) + lb;
            const long q0 = pair_class_off[cls], q1 = pair_class_off[cls + 1];
            if (q1 <= q0) continue;
            MDF_CHECK_ARG(q1 - q0 < (1L << 31), "mdf_int3c2e_ip: too many pairs in one class");
            for (int lc = 0; lc <= LMAX_AUX; lc++) {
                const long k0 = aux_l_off[lc], k1 = aux_l_off[lc + 1];
                if (k1 <= k0) continue;
                int rc = launch_int3c_ip(la, lb, lc, a, x, pairs + 2 * q0, (int)(q1 - q0), aux_sorted + k0,
                                         (int)(k1 - k0), W, ldw, col_map, ao_atom, aux_atom, natm, grad, grad_aux, st);
                if (rc) return rc;
            }
        }
    return MDF_OK;
}

// Nuclear derivative of the metric term: grad[3 * atom + x] +=
// sum_{P, Q} Z[P, Q] d(P|Q)/dR_atom,x, every ordered pair of aux shells
// (GPU4PySCF's int2c2e_ip1_per_atom, electrolites/fastdfgrad.py:1172-1178).
extern "C" int mdf_int2c2e_ip(const mdf_basis *aux, const int *aux_sorted, const long *aux_l_off,
                              const double *Z, long ldz, const int *aux_atom, int natm, double *grad,
                              void *stream) {
    MDF_CHECK_ARG(aux && aux_l_off && Z && aux_atom && grad, "mdf_int2c2e_ip: null argument");
    MDF_CHECK_ARG(natm > 0, "mdf_int2c2e_ip: natm must be positive");
    MDF_CHECK_ARG(aux->lmax <= LMAX_AUX, "mdf_int2c2e_ip: aux angular momentum %d > %d", aux->lmax, LMAX_AUX);
    hipStream_t st = as_stream(stream);
    const BasisDev x = to_dev(aux);
    for (int la = 0; la <= LMAX_AUX; la++) {
        const long p0 = aux_l_off[la], p1 = aux_l_off[la + 1];
        if (p1 <= p0) continue;
        for (int lc = 0; lc <= LMAX_AUX; lc++) {
            const long k0 = aux_l_off[lc], k1 = aux_l_off[lc + 1];
            if (k1 <= k0) continue;
            int rc = launch_int2c_ip(la, lc, x, aux_sorted + p0, (int)(p1 - p0), aux_sorted + k0,
                                     (int)(k1 - k0), Z, ldz, aux_atom, natm, grad, st);
            if (rc) return rc;
        }
    }
    return MDF_OK;
}

// Overlap, kinetic and nuclear-attraction matrices over the listed shell
// pairs (both triangles written; screened pairs untouched).  The one-electron
// part of the Fock matrix for the SCF driver (PySCF int1e_ovlp / int1e_kin /
// int1e_nuc).
extern "C" int mdf_int1e(const mdf_basis *ao, int nao, const int *pairs, const long *pair_class_off, int natm,
                         const double *charges, const double *atom_xyz, double *S, double *T, double *V,
                         void *stream) {
    MDF_CHECK_ARG(ao && pair_class_off && S && T && V && (natm == 0 || (charges && atom_xyz)),
                  "mdf_int1e: null argument");
    MDF_CHECK_ARG(ao->lmax <= LMAX_AO, "mdf_int1e: orbital l %d > %d", ao->lmax, LMAX_AO);
    hipStream_t st = as_stream(stream);
    const BasisDev a = to_dev(ao);
    MDF_CHECK_ARG(nao > 0, "mdf_int1e: nao must be positive");
    for (int la = 0; la <= LMAX_AO; la++)
        for (int lb = 0; lb <= LMAX_AO; lb++) {
            const int cls = la * (LMAX_AO + 1) + lb;
            const long q0 = pair_class_off[cls], q1 = pair_class_off[cls + 1];
            if (q1 <= q0) continue;
            MDF_CHECK_ARG(q1 - q0 < (1L << 31), "mdf_int1e: too many pairs in one class");
            int rc = launch_int1e(la, lb, a, pairs + 2 * q0, (int)(q1 - q0), natm, charges, atom_xyz, S, T, V,
                                  nao, st);
            if (rc) return rc;
        }
    return MDF_OK;
}
