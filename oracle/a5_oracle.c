/* CPU restatement of the Afivo tree core, ghost cells, restriction,
 * prolongation and FAS multigrid used by the a5_streamer hot path.
 * Test infrastructure only (see a5_oracle.h). Every routine cites the
 * reference routine it restates; operand order inside each expression is the
 * reference's, and the file is compiled with -ffp-contract=off so that no
 * multiply-add is fused.
 */
#include "a5_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NO_BOX 0
#define RM_REF (-1)
#define KP_REF 0
#define DO_REF 1
#define DEREFINE (-2)
#define REFINE 2
#define INIT_TAG (-2147483647) /* -huge(1), m_aX_t.f90:33 */
#define TAG_NORMAL 1
#define TAG_CEPS 3
#define TAG_VEPS 4

typedef struct {
  int n, cap;
  int *v;
} ivec;

static void iv_resize(ivec *a, int n) {
  if (n > a->cap) {
    a->cap = n + 16;
    a->v = (int *)realloc(a->v, sizeof(int) * (size_t)a->cap);
  }
  a->n = n;
}

struct otree {
  int dim, n_cell, n_cc, n_fc, coord, lvls_max, min_lvl, max_lvl, max_id;
  int cap; /* size(tree%boxes) */
  int ready;
  double dr_base, r_base[3];
  int nch, nnb, box_len, face_len;
  /* per box, index 1..cap */
  int *lvl, *ix, *parent, *children, *nbs, *in_use, *ncell, *tag;
  double *dr, *rmin;
  double *cc; /* [cap+1][n_cc][box_len] */
  double *fc; /* [cap+1][n_fc][dim][face_len] */
  ivec *ids, *leaves, *parents; /* index lvl - min_lvl, lvl in min..lvls_max+1 */
  /* ref_info (m_aX_t.f90:195-205) */
  int ri_n_lvls;
  ivec *ri_add, *ri_rm;
  /* multigrid options */
  int i_phi, i_tmp, i_rhs, i_eps, n_down, n_up, n_base, auto_ops, mg_ok;
  double helm; /* Helmholtz coefficient kappa^2 of the operator (0: Laplacian) */
  int bc_kind[6];
  double bc_val[6];
};

/* ---- topology tables (m_aX_t.f90:41-132), as closed forms ---- */
static int ch_dix(int c, int d) { return ((c - 1) >> d) & 1; }
static int ch_rev(int c, int d) { return ((c - 1) ^ (1 << d)) + 1; }
static int nb_dim(int nb) { return (nb - 1) / 2; } /* 0-based */
static int nb_low(int nb) { return ((nb - 1) & 1) == 0; }
static int nb_hi01(int nb) { return (nb - 1) & 1; }
static int nb_rev(int nb) { return ((nb - 1) ^ 1) + 1; }
static int nb_dix(int nb, int d) {
  if (nb_dim(nb) != d) return 0;
  return nb_low(nb) ? -1 : 1;
}
/* the n-th (0-based) child adjacent to neighbour direction nb */
static int ch_adj_nb(const otree *t, int n, int nb) {
  int d = nb_dim(nb), want = nb_hi01(nb), cnt = 0;
  for (int c = 1; c <= t->nch; c++)
    if (ch_dix(c, d) == want) {
      if (cnt == n) return c;
      cnt++;
    }
  return -1;
}

#define LV(t, l) ((l) - (t)->min_lvl)
#define IXP(t, id) (&(t)->ix[(size_t)(id) * 3])
#define CHP(t, id) (&(t)->children[(size_t)(id) * 8])
#define NBP(t, id) (&(t)->nbs[(size_t)(id) * 6])
#define RMP(t, id) (&(t)->rmin[(size_t)(id) * 3])

static double *CC(const otree *t, int id, int iv) {
  return t->cc + ((size_t)id * t->n_cc + (size_t)(iv - 1)) * t->box_len;
}
static double *FC(const otree *t, int id, int ifc, int d) {
  return t->fc + (((size_t)id * t->n_fc + (size_t)(ifc - 1)) * t->dim + d) *
                     t->face_len;
}
/* cell (i,j,k) of a box with nc cells per dimension, ghosts included */
static inline size_t I2(int nc, int i, int j) { return (size_t)i + (size_t)(nc + 2) * j; }
static inline size_t I3(int nc, int i, int j, int k) {
  return (size_t)i + (size_t)(nc + 2) * ((size_t)j + (size_t)(nc + 2) * k);
}
/* face arrays fx(nc+1,nc[,nc]), fy(nc,nc+1[,nc]), fz(nc,nc,nc+1), 1-based */
static inline size_t FI(int nc, int dim, int d, int i, int j, int k) {
  int n0 = nc + (d == 0), n1 = nc + (d == 1), n2 = nc + (d == 2);
  (void)n2;
  if (dim == 2) return (size_t)(i - 1) + (size_t)n0 * (j - 1);
  return (size_t)(i - 1) + (size_t)n0 * ((size_t)(j - 1) + (size_t)n1 * (k - 1));
}

/* ------------------------------------------------------------------ */
/* tree storage                                                         */
/* ------------------------------------------------------------------ */

static void resize_storage(otree *t, int new_cap) {
  /* a$D_resize_box_storage, m_aX_core.f90:448-464 */
  int old = t->cap;
  t->lvl = realloc(t->lvl, sizeof(int) * (size_t)(new_cap + 1));
  t->ix = realloc(t->ix, sizeof(int) * 3 * (size_t)(new_cap + 1));
  t->parent = realloc(t->parent, sizeof(int) * (size_t)(new_cap + 1));
  t->children = realloc(t->children, sizeof(int) * 8 * (size_t)(new_cap + 1));
  t->nbs = realloc(t->nbs, sizeof(int) * 6 * (size_t)(new_cap + 1));
  t->in_use = realloc(t->in_use, sizeof(int) * (size_t)(new_cap + 1));
  t->ncell = realloc(t->ncell, sizeof(int) * (size_t)(new_cap + 1));
  t->tag = realloc(t->tag, sizeof(int) * (size_t)(new_cap + 1));
  t->dr = realloc(t->dr, sizeof(double) * (size_t)(new_cap + 1));
  t->rmin = realloc(t->rmin, sizeof(double) * 3 * (size_t)(new_cap + 1));
  t->cc = realloc(t->cc, sizeof(double) * (size_t)(new_cap + 1) * t->n_cc * t->box_len);
  t->fc = realloc(t->fc, sizeof(double) * (size_t)(new_cap + 1) *
                             (t->n_fc > 0 ? t->n_fc : 1) * t->dim * t->face_len);
  for (int id = old + 1; id <= new_cap; id++) {
    t->in_use[id] = 0;
    t->tag[id] = INIT_TAG;
    t->lvl[id] = 0;
    t->parent[id] = 0;
    memset(IXP(t, id), 0, sizeof(int) * 3);
    memset(CHP(t, id), 0, sizeof(int) * 8);
    memset(NBP(t, id), 0, sizeof(int) * 6);
    memset(CC(t, id, 1), 0, sizeof(double) * t->n_cc * t->box_len);
    memset(t->fc + (size_t)id * (t->n_fc > 0 ? t->n_fc : 1) * t->dim * t->face_len, 0,
           sizeof(double) * (t->n_fc > 0 ? t->n_fc : 1) * t->dim * t->face_len);
  }
  if (new_cap < old) {
    /* shrinking is only used by tidy_up, which rewrites all boxes */
  }
  t->cap = new_cap;
}

otree *o_tree_new(int dim, int n_cell, int n_cc, int n_fc, double dr,
                  const double *r_min, int coarsen_to, int lvls_max, int coord,
                  int n_boxes) {
  /* a$D_init, m_aX_core.f90:47-116 */
  otree *t = (otree *)calloc(1, sizeof(otree));
  t->dim = dim;
  t->n_cell = n_cell;
  t->n_cc = n_cc;
  t->n_fc = n_fc;
  t->coord = coord;
  t->lvls_max = lvls_max > 0 ? lvls_max : 30;
  t->dr_base = dr;
  for (int d = 0; d < dim; d++) t->r_base[d] = r_min ? r_min[d] : 0.0;
  t->nch = 1 << dim;
  t->nnb = 2 * dim;
  t->box_len = (dim == 2) ? (n_cell + 2) * (n_cell + 2)
                          : (n_cell + 2) * (n_cell + 2) * (n_cell + 2);
  t->face_len = (dim == 2) ? (n_cell + 1) * n_cell : (n_cell + 1) * n_cell * n_cell;
  if (coarsen_to > 0) {
    t->min_lvl = 1 - (int)lround(log((double)n_cell / coarsen_to) / log(2.0));
  } else {
    t->min_lvl = 1;
  }
  int nl = t->lvls_max + 1 - t->min_lvl + 1;
  t->ids = calloc((size_t)nl, sizeof(ivec));
  t->leaves = calloc((size_t)nl, sizeof(ivec));
  t->parents = calloc((size_t)nl, sizeof(ivec));
  t->cap = 0;
  resize_storage(t, n_boxes > 0 ? n_boxes : 100);
  t->max_id = 0;
  t->max_lvl = 0;
  t->i_eps = -1;
  return t;
}

void o_tree_free(otree *t) {
  if (!t) return;
  int nl = t->lvls_max + 1 - t->min_lvl + 1;
  for (int l = 0; l < nl; l++) {
    free(t->ids[l].v);
    free(t->leaves[l].v);
    free(t->parents[l].v);
  }
  for (int l = 0; l < t->ri_n_lvls; l++) {
    free(t->ri_add[l].v);
    free(t->ri_rm[l].v);
  }
  free(t->ri_add);
  free(t->ri_rm);
  free(t->ids);
  free(t->leaves);
  free(t->parents);
  free(t->lvl); free(t->ix); free(t->parent); free(t->children); free(t->nbs);
  free(t->in_use); free(t->ncell); free(t->tag); free(t->dr); free(t->rmin);
  free(t->cc); free(t->fc);
  free(t);
}

static int has_children(const otree *t, int id) { return CHP(t, id)[0] != NO_BOX; }

static void get_free_ids(otree *t, int n, int *out) {
  /* get_free_ids, m_aX_core.f90:605-617 */
  for (int i = 0; i < n; i++) out[i] = t->max_id + 1 + i;
  t->max_id += n;
}

static void init_box(otree *t, int id) {
  /* init_box, m_aX_core.f90:363-381 (storage is preallocated) */
  t->in_use[id] = 1;
}

static void set_leaves_parents(otree *t, int lvl) {
  /* set_leaves_parents, m_aX_core.f90:328-359 */
  ivec *ids = &t->ids[LV(t, lvl)], *lv = &t->leaves[LV(t, lvl)],
       *pa = &t->parents[LV(t, lvl)];
  int np = 0;
  for (int i = 0; i < ids->n; i++) np += has_children(t, ids->v[i]);
  iv_resize(pa, np);
  iv_resize(lv, ids->n - np);
  int il = 0, ip = 0;
  for (int i = 0; i < ids->n; i++) {
    int id = ids->v[i];
    if (has_children(t, id)) pa->v[ip++] = id;
    else lv->v[il++] = id;
  }
}

void o_set_base(otree *t, int n, const int *ix_list, const int *nb_list_in) {
  /* a$D_set_base, m_aX_core.f90:137-220 */
  int nnb = t->nnb, D = t->dim;
  int *nbl = malloc(sizeof(int) * (size_t)nnb * n);
  memcpy(nbl, nb_list_in, sizeof(int) * (size_t)nnb * n);
  for (int i = 1; i <= n; i++)
    for (int nb = 1; nb <= nnb; nb++) {
      int nb_id = nbl[(i - 1) * nnb + nb - 1];
      if (nb_id > NO_BOX && nb_id != i) nbl[(nb_id - 1) * nnb + nb_rev(nb) - 1] = i;
    }
  if (n > t->cap) resize_storage(t, n);
  for (int lvl = t->min_lvl; lvl <= 1; lvl++) {
    ivec *ids = &t->ids[LV(t, lvl)];
    iv_resize(ids, n);
    get_free_ids(t, n, ids->v);
    int offset = ids->v[0] - 1;
    for (int i = 1; i <= n; i++) {
      int id = ids->v[i - 1];
      if (id > t->cap) resize_storage(t, id + 100);
      t->lvl[id] = lvl;
      for (int d = 0; d < D; d++) {
        IXP(t, id)[d] = ix_list[(i - 1) * D + d];
        RMP(t, id)[d] = t->r_base[d] + (ix_list[(i - 1) * D + d] - 1) * t->dr_base * t->n_cell;
      }
      t->dr[id] = t->dr_base * pow(0.5, lvl - 1);
      t->ncell[id] = t->n_cell / (1 << (1 - lvl));
      t->parent[id] = NO_BOX;
      for (int c = 0; c < 8; c++) CHP(t, id)[c] = NO_BOX;
      for (int nb = 1; nb <= nnb; nb++) {
        int v = nbl[(i - 1) * nnb + nb - 1];
        NBP(t, id)[nb - 1] = v > NO_BOX ? v + offset : v;
      }
      init_box(t, id);
    }
    if (lvl == 1) {
      iv_resize(&t->leaves[LV(t, lvl)], n);
      memcpy(t->leaves[LV(t, lvl)].v, ids->v, sizeof(int) * (size_t)n);
    } else {
      iv_resize(&t->parents[LV(t, lvl)], n);
      memcpy(t->parents[LV(t, lvl)].v, ids->v, sizeof(int) * (size_t)n);
    }
    if (lvl > t->min_lvl) {
      ivec *pids = &t->ids[LV(t, lvl - 1)];
      for (int i = 0; i < n; i++) {
        CHP(t, pids->v[i])[0] = ids->v[i];
        t->parent[ids->v[i]] = pids->v[i];
      }
    }
  }
  free(nbl);
  t->max_lvl = 1;
  t->ready = 1;
}

static int ix_to_cix(const otree *t, int id) {
  /* a$D_ix_to_cix, m_aX_core.f90:416-424 */
  const int *ix = IXP(t, id);
  if (t->dim == 2) return 4 - 2 * (ix[1] & 1) - (ix[0] & 1);
  return 8 - 4 * (ix[2] & 1) - 2 * (ix[1] & 1) - (ix[0] & 1);
}

static int find_nb(const otree *t, int id, int nb) {
  /* find_nb_$Dd, m_aX_core.f90:427-445 */
  int p_id = t->parent[id];
  int c_ix = ix_to_cix(t, id);
  int d = nb_dim(nb);
  int low = ch_dix(c_ix, d) == 0;
  if (low == nb_low(nb)) p_id = NBP(t, p_id)[nb - 1];
  return CHP(t, p_id)[ch_rev(c_ix, d) - 1];
}

static void set_nbs(otree *t, int id) {
  /* set_nbs_$Dd, m_aX_core.f90:399-413 */
  for (int nb = 1; nb <= t->nnb; nb++) {
    if (NBP(t, id)[nb - 1] == NO_BOX) {
      int nb_id = find_nb(t, id, nb);
      if (nb_id > NO_BOX) {
        NBP(t, id)[nb - 1] = nb_id;
        NBP(t, nb_id)[nb_rev(nb) - 1] = id;
      }
    }
  }
}

static void add_children(otree *t, int id, const int *c_ids) {
  /* add_children, m_aX_core.f90:753-788 */
  int D = t->dim;
  for (int i = 1; i <= t->nch; i++) CHP(t, id)[i - 1] = c_ids[i - 1];
  for (int i = 1; i <= t->nch; i++) {
    int c = c_ids[i - 1];
    for (int d = 0; d < D; d++) {
      IXP(t, c)[d] = 2 * IXP(t, id)[d] - 1 + ch_dix(i, d);
      RMP(t, c)[d] = RMP(t, id)[d] + 0.5 * t->dr[id] * ch_dix(i, d) * t->ncell[id];
    }
    t->lvl[c] = t->lvl[id] + 1;
    t->parent[c] = id;
    t->tag[c] = INIT_TAG;
    for (int k = 0; k < 8; k++) CHP(t, c)[k] = NO_BOX;
    for (int k = 0; k < 6; k++) NBP(t, c)[k] = NO_BOX;
    t->ncell[c] = t->ncell[id];
    t->dr[c] = 0.5 * t->dr[id];
    init_box(t, c);
  }
  for (int nb = 1; nb <= t->nnb; nb++) {
    if (NBP(t, id)[nb - 1] < NO_BOX) {
      for (int n = 0; n < t->nch / 2; n++) {
        int c = c_ids[ch_adj_nb(t, n, nb) - 1];
        NBP(t, c)[nb - 1] = NBP(t, id)[nb - 1];
      }
    }
  }
}

static void remove_children(otree *t, int id) {
  /* remove_children, m_aX_core.f90:729-750 */
  for (int ic = 1; ic <= t->nch; ic++) {
    int c = CHP(t, id)[ic - 1];
    for (int nb = 1; nb <= t->nnb; nb++) {
      int nb_id = NBP(t, c)[nb - 1];
      if (nb_id > NO_BOX) NBP(t, nb_id)[nb_rev(nb) - 1] = NO_BOX;
    }
    t->in_use[c] = 0;
  }
  for (int ic = 0; ic < 8; ic++) CHP(t, id)[ic] = NO_BOX;
}

static void set_child_ids(otree *t, int lvl) {
  /* set_child_ids, m_aX_core.f90:792-809 */
  ivec *pa = &t->parents[LV(t, lvl)], *c = &t->ids[LV(t, lvl + 1)];
  iv_resize(c, t->nch * pa->n);
  for (int i = 0; i < pa->n; i++)
    for (int k = 0; k < t->nch; k++) c->v[i * t->nch + k] = CHP(t, pa->v[i])[k];
}

static void consistent_ref_flags(otree *t, int *f, int n_in, const int *flags_in) {
  /* consistent_ref_flags, m_aX_core.f90:624-726; the user callback is the
   * flag array given by the caller, read for boxes of levels 1..max_lvl */
  int n = t->max_id;
  for (int i = 1; i <= n; i++) f[i] = -2147483647;
  for (int lvl = 1; lvl <= t->max_lvl; lvl++) {
    ivec *ids = &t->ids[LV(t, lvl)];
    for (int i = 0; i < ids->n; i++) {
      int id = ids->v[i];
      f[id] = (id <= n_in) ? flags_in[id - 1] : KP_REF;
    }
  }
  for (int i = 1; i <= n; i++) {
    if (f[i] > DO_REF) f[i] = KP_REF;
    if (f[i] < RM_REF) f[i] = KP_REF;
  }
  ivec *top = &t->ids[LV(t, t->lvls_max)];
  for (int i = 0; i < top->n; i++)
    if (f[top->v[i]] == DO_REF) f[top->v[i]] = KP_REF;

  for (int lvl = t->max_lvl; lvl >= 1; lvl--) {
    ivec *lv = &t->leaves[LV(t, lvl)];
    for (int i = 0; i < lv->n; i++) {
      int id = lv->v[i];
      if (f[id] > KP_REF) {
        f[id] = REFINE;
        for (int nb = 1; nb <= t->nnb; nb++) {
          int nb_id = NBP(t, id)[nb - 1];
          if (nb_id == NO_BOX) {
            int p_id = t->parent[id];
            int p_nb_id = NBP(t, p_id)[nb - 1];
            f[p_nb_id] = REFINE;
          }
        }
      } else if (f[id] == RM_REF) {
        for (int nb = 1; nb <= t->nnb; nb++) {
          int nb_id = NBP(t, id)[nb - 1];
          if (nb_id > NO_BOX) {
            if (has_children(t, nb_id) || f[nb_id] > KP_REF) {
              f[id] = KP_REF;
              break;
            }
          }
        }
      }
    }
  }
  for (int lvl = t->max_lvl - 1; lvl >= 1; lvl--) {
    ivec *pa = &t->parents[LV(t, lvl)];
    for (int i = 0; i < pa->n; i++) {
      int id = pa->v[i], all_rm = 1;
      for (int c = 0; c < t->nch; c++)
        if (f[CHP(t, id)[c]] != RM_REF) all_rm = 0;
      if (all_rm && f[id] <= KP_REF) f[id] = DEREFINE;
      else f[id] = KP_REF;
    }
  }
}

static void set_ref_info(otree *t, const int *f, int n) {
  /* set_ref_info, m_aX_core.f90:540-601 */
  for (int l = 0; l < t->ri_n_lvls; l++) {
    free(t->ri_add[l].v);
    free(t->ri_rm[l].v);
  }
  free(t->ri_add);
  free(t->ri_rm);
  t->ri_n_lvls = t->max_lvl + 1; /* levels 1..max_lvl+1 */
  t->ri_add = calloc((size_t)t->ri_n_lvls + 1, sizeof(ivec));
  t->ri_rm = calloc((size_t)t->ri_n_lvls + 1, sizeof(ivec));
  for (int id = 1; id <= n; id++) {
    int lvl = t->lvl[id];
    if (lvl < 1 || lvl + 1 > t->ri_n_lvls) continue;
    if (f[id] == REFINE) {
      ivec *a = &t->ri_add[lvl + 1 - 1];
      int o = a->n;
      iv_resize(a, o + t->nch);
      for (int c = 0; c < t->nch; c++) a->v[o + c] = CHP(t, id)[c];
    } else if (f[id] == DEREFINE) {
      ivec *a = &t->ri_rm[lvl + 1 - 1];
      int o = a->n;
      iv_resize(a, o + t->nch);
      for (int c = 0; c < t->nch; c++) a->v[o + c] = CHP(t, id)[c];
    }
  }
}

void o_adjust(otree *t, int n_flags, const int *flags, int *n_add, int *n_rm) {
  /* a$D_adjust_refinement, m_aX_core.f90:473-537 */
  int max_id_prev = t->max_id;
  int *f = malloc(sizeof(int) * (size_t)(max_id_prev + 1));
  consistent_ref_flags(t, f, n_flags, flags);
  int nref = 0, nder = 0;
  for (int i = 1; i <= max_id_prev; i++) {
    nref += f[i] == REFINE;
    nder += f[i] == DEREFINE;
  }
  int max_id_req = max_id_prev + t->nch * nref;
  if (max_id_req > t->cap) resize_storage(t, max_id_req);
  int lvl;
  for (lvl = 1; lvl <= t->lvls_max - 1; lvl++) {
    ivec *ids = &t->ids[LV(t, lvl)];
    for (int i = 0; i < ids->n; i++) {
      int id = ids->v[i];
      if (id > max_id_prev) continue;
      if (f[id] == REFINE) {
        int c_ids[8];
        get_free_ids(t, t->nch, c_ids);
        add_children(t, id, c_ids);
      } else if (f[id] == DEREFINE) {
        remove_children(t, id);
      }
    }
    set_leaves_parents(t, lvl);
    set_child_ids(t, lvl);
    ivec *nx = &t->ids[LV(t, lvl + 1)];
    for (int i = 0; i < nx->n; i++)
      if (nx->v[i] > max_id_prev) set_nbs(t, nx->v[i]);
    if (nx->n == 0) break;
  }
  t->max_lvl = lvl;
  set_leaves_parents(t, t->max_lvl + 1);
  set_ref_info(t, f, max_id_prev);
  *n_add = t->nch * nref;
  *n_rm = t->nch * nder;
  free(f);
}

int o_ref_info_lvl(const otree *t, int lvl, int *add, int *rm, int *n_add,
                   int *n_rm) {
  *n_add = 0;
  *n_rm = 0;
  if (lvl < 1 || lvl > t->ri_n_lvls) return 0;
  *n_add = t->ri_add[lvl - 1].n;
  *n_rm = t->ri_rm[lvl - 1].n;
  if (add) memcpy(add, t->ri_add[lvl - 1].v, sizeof(int) * (size_t)*n_add);
  if (rm) memcpy(rm, t->ri_rm[lvl - 1].v, sizeof(int) * (size_t)*n_rm);
  return 0;
}

/* Morton keys, m_morton.f90:33-59 (bit interleave, x in the lowest bit) */
static uint64_t morton_key(const otree *t, int id) {
  uint64_t m = 0;
  int D = t->dim;
  for (int b = 0; b < 21; b++)
    for (int d = 0; d < D; d++) {
      uint64_t v = (uint64_t)(IXP(t, id)[d] - 1);
      m |= ((v >> b) & 1ull) << (b * D + d);
    }
  return m;
}

typedef struct {
  uint64_t key;
  int pos;
} mkey;

static int mkey_cmp(const void *a, const void *b) {
  const mkey *x = a, *y = b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return x->pos - y->pos; /* stable, as morton_rank (m_morton.f90:155-353) */
}

void o_tidy_up(otree *t, double max_frac, double goal_frac, int n_clean_min,
               int reorder) {
  /* a$D_tidy_up, m_aX_core.f90:224-325 */
  int max_id = t->max_id, n_used = 0;
  for (int i = 1; i <= max_id; i++) n_used += t->in_use[i];
  int old_size = t->cap;
  double frac_in_use = n_used / (double)old_size;
  int n_clean = (int)lround((goal_frac - frac_in_use) * old_size);
  int new_size = old_size;
  if (!reorder) {
    if (max_id > old_size * max_frac ||
        (frac_in_use < goal_frac && n_clean > n_clean_min)) {
      long v = lround(n_used / goal_frac);
      new_size = v > 1 ? (int)v : 1;
    }
  }
  if (new_size == old_size && !reorder) return;

  int ncap = reorder ? (n_used > old_size ? n_used : old_size) : new_size;
  if (ncap < n_used) ncap = n_used;
  int *map = calloc((size_t)max_id + 1, sizeof(int));
  int *order = malloc(sizeof(int) * (size_t)(n_used + 1));
  int n_stored = 0;
  for (int lvl = t->min_lvl; lvl <= t->max_lvl; lvl++) {
    ivec *ids = &t->ids[LV(t, lvl)];
    int n = ids->n;
    mkey *k = malloc(sizeof(mkey) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; i++) {
      k[i].key = morton_key(t, ids->v[i]);
      k[i].pos = i;
    }
    qsort(k, (size_t)n, sizeof(mkey), mkey_cmp);
    int *sorted = malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; i++) sorted[i] = ids->v[k[i].pos];
    for (int i = 0; i < n; i++) {
      order[n_stored + i + 1] = sorted[i];
      map[sorted[i]] = n_stored + i + 1;
    }
    for (int i = 0; i < n; i++) ids->v[i] = n_stored + i + 1;
    n_stored += n;
    free(k);
    free(sorted);
  }
  /* gather box records into the new order */
  otree tmp = *t;
  tmp.lvl = malloc(sizeof(int) * (size_t)(ncap + 1));
  tmp.ix = calloc((size_t)(ncap + 1) * 3, sizeof(int));
  tmp.parent = malloc(sizeof(int) * (size_t)(ncap + 1));
  tmp.children = calloc((size_t)(ncap + 1) * 8, sizeof(int));
  tmp.nbs = calloc((size_t)(ncap + 1) * 6, sizeof(int));
  tmp.in_use = calloc((size_t)(ncap + 1), sizeof(int));
  tmp.ncell = calloc((size_t)(ncap + 1), sizeof(int));
  tmp.tag = malloc(sizeof(int) * (size_t)(ncap + 1));
  tmp.dr = calloc((size_t)(ncap + 1), sizeof(double));
  tmp.rmin = calloc((size_t)(ncap + 1) * 3, sizeof(double));
  tmp.cc = calloc((size_t)(ncap + 1) * t->n_cc * t->box_len, sizeof(double));
  size_t fcl = (size_t)(t->n_fc > 0 ? t->n_fc : 1) * t->dim * t->face_len;
  tmp.fc = calloc((size_t)(ncap + 1) * fcl, sizeof(double));
  for (int i = 0; i <= ncap; i++) tmp.tag[i] = INIT_TAG, tmp.lvl[i] = 0, tmp.parent[i] = 0;
  for (int n = 1; n <= n_used; n++) {
    int o = order[n];
    tmp.lvl[n] = t->lvl[o];
    memcpy(&tmp.ix[n * 3], IXP(t, o), sizeof(int) * 3);
    tmp.parent[n] = map[t->parent[o]];
    for (int c = 0; c < 8; c++) tmp.children[n * 8 + c] = map[CHP(t, o)[c]];
    for (int b = 0; b < 6; b++) {
      int v = NBP(t, o)[b];
      tmp.nbs[n * 6 + b] = v > NO_BOX ? map[v] : v;
    }
    tmp.in_use[n] = 1;
    tmp.ncell[n] = t->ncell[o];
    tmp.tag[n] = t->tag[o];
    tmp.dr[n] = t->dr[o];
    memcpy(&tmp.rmin[n * 3], RMP(t, o), sizeof(double) * 3);
    memcpy(tmp.cc + (size_t)n * t->n_cc * t->box_len, CC(t, o, 1),
           sizeof(double) * t->n_cc * t->box_len);
    memcpy(tmp.fc + (size_t)n * fcl, t->fc + (size_t)o * fcl, sizeof(double) * fcl);
  }
  free(t->lvl); free(t->ix); free(t->parent); free(t->children); free(t->nbs);
  free(t->in_use); free(t->ncell); free(t->tag); free(t->dr); free(t->rmin);
  free(t->cc); free(t->fc);
  t->lvl = tmp.lvl; t->ix = tmp.ix; t->parent = tmp.parent;
  t->children = tmp.children; t->nbs = tmp.nbs; t->in_use = tmp.in_use;
  t->ncell = tmp.ncell; t->tag = tmp.tag; t->dr = tmp.dr; t->rmin = tmp.rmin;
  t->cc = tmp.cc; t->fc = tmp.fc;
  t->cap = reorder ? old_size : new_size;
  if (t->cap < n_used) t->cap = n_used;
  if (ncap > t->cap) t->cap = ncap;
  for (int lvl = t->min_lvl; lvl <= t->max_lvl; lvl++) set_leaves_parents(t, lvl);
  t->max_id = n_used;
  free(map);
  free(order);
}

void o_info(const otree *t, int *out) {
  out[0] = t->max_id;
  out[1] = t->max_lvl;
  out[2] = t->min_lvl;
  out[3] = t->lvls_max;
  out[4] = t->cap;
  out[5] = t->n_cell;
  out[6] = t->coord;
  out[7] = t->ready;
}

void o_get_boxes(const otree *t, int n, int *lvl, int *ix, int *parent,
                 int *children, int *neighbors, int *in_use, int *n_cell,
                 double *dr, double *r_min) {
  int D = t->dim, nch = t->nch, nnb = t->nnb;
  for (int id = 1; id <= n; id++) {
    int u = id <= t->cap && t->in_use[id];
    in_use[id - 1] = u;
    lvl[id - 1] = u ? t->lvl[id] : 0;
    parent[id - 1] = u ? t->parent[id] : 0;
    n_cell[id - 1] = u ? t->ncell[id] : 0;
    dr[id - 1] = u ? t->dr[id] : 0;
    for (int d = 0; d < D; d++) {
      ix[(id - 1) * D + d] = u ? IXP(t, id)[d] : 0;
      r_min[(id - 1) * D + d] = u ? RMP(t, id)[d] : 0;
    }
    for (int c = 0; c < nch; c++) children[(id - 1) * nch + c] = u ? CHP(t, id)[c] : 0;
    for (int b = 0; b < nnb; b++) neighbors[(id - 1) * nnb + b] = u ? NBP(t, id)[b] : 0;
  }
}

void o_level_counts(const otree *t, int lvl, int *c) {
  c[0] = c[1] = c[2] = 0;
  if (lvl < t->min_lvl || lvl > t->lvls_max + 1) return;
  c[0] = t->ids[LV(t, lvl)].n;
  c[1] = t->leaves[LV(t, lvl)].n;
  c[2] = t->parents[LV(t, lvl)].n;
}

void o_level_lists(const otree *t, int lvl, int *ids, int *leaves, int *parents) {
  memcpy(ids, t->ids[LV(t, lvl)].v, sizeof(int) * (size_t)t->ids[LV(t, lvl)].n);
  memcpy(leaves, t->leaves[LV(t, lvl)].v, sizeof(int) * (size_t)t->leaves[LV(t, lvl)].n);
  memcpy(parents, t->parents[LV(t, lvl)].v, sizeof(int) * (size_t)t->parents[LV(t, lvl)].n);
}

static int blen(const otree *t, int id) {
  int nc = t->ncell[id];
  return t->dim == 2 ? (nc + 2) * (nc + 2) : (nc + 2) * (nc + 2) * (nc + 2);
}

void o_set_cc(otree *t, int id, int iv, const double *data) {
  memcpy(CC(t, id, iv), data, sizeof(double) * (size_t)blen(t, id));
}
void o_get_cc(const otree *t, int id, int iv, double *data) {
  memcpy(data, CC(t, id, iv), sizeof(double) * (size_t)blen(t, id));
}
void o_set_fc(otree *t, int id, int dim, int ifc, const double *data) {
  memcpy(FC(t, id, ifc, dim - 1), data, sizeof(double) * (size_t)t->face_len);
}
void o_get_fc(const otree *t, int id, int dim, int ifc, double *data) {
  memcpy(data, FC(t, id, ifc, dim - 1), sizeof(double) * (size_t)t->face_len);
}

/* ------------------------------------------------------------------ */
/* utilities (m_aX_utils.f90)                                           */
/* ------------------------------------------------------------------ */

static void child_offset(const otree *t, int id, int nb, int *off) {
  /* a$D_get_child_offset, m_aX_utils.f90:227-239 */
  int nc = t->ncell[id];
  for (int d = 0; d < t->dim; d++) {
    if (t->lvl[id] > 1) {
      off[d] = ((IXP(t, id)[d] - 1) & 1) * (nc >> 1);
      if (nb > 0) off[d] -= nb_dix(nb, d) * nc;
    } else {
      off[d] = 0;
      if (nb > 0) off[d] -= nb_dix(nb, d) * (nc >> 1);
    }
  }
}

static void box_copy_cc(otree *t, int id, int from, int to) {
  /* a$D_box_copy_cc, m_aX_utils.f90:421-429 (whole box, ghosts included) */
  memcpy(CC(t, id, to), CC(t, id, from), sizeof(double) * (size_t)blen(t, id));
}

static void box_add_cc(otree *t, int id, int from, int to) {
  /* a$D_box_add_cc, m_aX_utils.f90:363-371 */
  double *a = CC(t, id, to), *b = CC(t, id, from);
  int n = blen(t, id);
  for (int i = 0; i < n; i++) a[i] = a[i] + b[i];
}

static void box_clear_cc(otree *t, int id, int iv) {
  /* a$D_box_clear_cc, m_aX_utils.f90:352-360 */
  double *a = CC(t, id, iv);
  int n = blen(t, id);
  for (int i = 0; i < n; i++) a[i] = 0;
}

/* ------------------------------------------------------------------ */
/* prolongation (m_aX_prolong.f90)                                      */
/* ------------------------------------------------------------------ */

static void prolong0_to(otree *t, int id, int iv, const int *lo, const int *hi) {
  /* a$D_prolong0_to, m_aX_prolong.f90:37-78 */
  int nc = t->ncell[id], p = t->parent[id], pnc = t->ncell[p], off[3];
  child_offset(t, id, 0, off);
  double *c = CC(t, id, iv), *pc = CC(t, p, iv);
  if (t->dim == 2) {
    for (int j = lo[1]; j <= hi[1]; j++) {
      int jc = off[1] + ((j + 1) >> 1);
      for (int i = lo[0]; i <= hi[0]; i++) {
        int ic = off[0] + ((i + 1) >> 1);
        c[I2(nc, i, j)] = pc[I2(pnc, ic, jc)];
      }
    }
  } else {
    for (int k = lo[2]; k <= hi[2]; k++) {
      int kc = off[2] + ((k + 1) >> 1);
      for (int j = lo[1]; j <= hi[1]; j++) {
        int jc = off[1] + ((j + 1) >> 1);
        for (int i = lo[0]; i <= hi[0]; i++) {
          int ic = off[0] + ((i + 1) >> 1);
          c[I3(nc, i, j, k)] = pc[I3(pnc, ic, jc, kc)];
        }
      }
    }
  }
}

void o_prolong0_to(otree *t, int id, int iv) {
  int nc = t->ncell[id], lo[3] = {1, 1, 1}, hi[3] = {nc, nc, nc};
  prolong0_to(t, id, iv, lo, hi);
}

void o_prolong1_to(otree *t, int id, int iv, int add) {
  /* a$D_prolong1_to, m_aX_prolong.f90:99-191 */
  int nc = t->ncell[id], hnc = nc >> 1, p = t->parent[id], pnc = t->ncell[p], off[3];
  child_offset(t, id, 0, off);
  double *c = CC(t, id, iv), *pc = CC(t, p, iv);
  if (t->dim == 2) {
    if (!add)
      for (int j = 1; j <= nc; j++)
        for (int i = 1; i <= nc; i++) c[I2(nc, i, j)] = 0;
    for (int j = 1; j <= hnc; j++) {
      int jc = j + off[1], jf = 2 * j - 1;
      for (int i = 1; i <= hnc; i++) {
        int ic = i + off[0], fi = 2 * i - 1;
        double f0 = 0.5 * pc[I2(pnc, ic, jc)];
        double flx = 0.25 * pc[I2(pnc, ic - 1, jc)];
        double fhx = 0.25 * pc[I2(pnc, ic + 1, jc)];
        double fly = 0.25 * pc[I2(pnc, ic, jc - 1)];
        double fhy = 0.25 * pc[I2(pnc, ic, jc + 1)];
        c[I2(nc, fi, jf)] = f0 + flx + fly + c[I2(nc, fi, jf)];
        c[I2(nc, fi + 1, jf)] = f0 + fhx + fly + c[I2(nc, fi + 1, jf)];
        c[I2(nc, fi, jf + 1)] = f0 + flx + fhy + c[I2(nc, fi, jf + 1)];
        c[I2(nc, fi + 1, jf + 1)] = f0 + fhx + fhy + c[I2(nc, fi + 1, jf + 1)];
      }
    }
  } else {
    if (!add)
      for (int k = 1; k <= nc; k++)
        for (int j = 1; j <= nc; j++)
          for (int i = 1; i <= nc; i++) c[I3(nc, i, j, k)] = 0;
    for (int k = 1; k <= hnc; k++) {
      int kc = k + off[2], kf = 2 * k - 1;
      for (int j = 1; j <= hnc; j++) {
        int jc = j + off[1], jf = 2 * j - 1;
        for (int i = 1; i <= hnc; i++) {
          int ic = i + off[0], fi = 2 * i - 1;
          double f0 = 0.25 * pc[I3(pnc, ic, jc, kc)];
          double flx = 0.25 * pc[I3(pnc, ic - 1, jc, kc)];
          double fhx = 0.25 * pc[I3(pnc, ic + 1, jc, kc)];
          double fly = 0.25 * pc[I3(pnc, ic, jc - 1, kc)];
          double fhy = 0.25 * pc[I3(pnc, ic, jc + 1, kc)];
          double flz = 0.25 * pc[I3(pnc, ic, jc, kc - 1)];
          double fhz = 0.25 * pc[I3(pnc, ic, jc, kc + 1)];
#define P1(di, dj, dk, fx, fy, fz) \
  c[I3(nc, fi + di, jf + dj, kf + dk)] = f0 + fx + fy + fz + c[I3(nc, fi + di, jf + dj, kf + dk)]
          P1(0, 0, 0, flx, fly, flz);
          P1(1, 0, 0, fhx, fly, flz);
          P1(0, 1, 0, flx, fhy, flz);
          P1(1, 1, 0, fhx, fhy, flz);
          P1(0, 0, 1, flx, fly, fhz);
          P1(1, 0, 1, fhx, fly, fhz);
          P1(0, 1, 1, flx, fhy, fhz);
          P1(1, 1, 1, fhx, fhy, fhz);
#undef P1
        }
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* ghost cells (m_aX_gc.f90)                                            */
/* ------------------------------------------------------------------ */

/* Visit the face cells of box id next to neighbour nb: g = ghost index,
 * c = first interior cell, c2 = second interior cell (flat indices), and
 * the tangential cell indices (a, b) in 1..nc. */
#define FACE_LOOP_BEGIN(t, id, nb)                                        \
  {                                                                       \
    int nc_ = t->ncell[id], D_ = t->dim, d_ = nb_dim(nb);                 \
    int gix_ = nb_low(nb) ? 0 : nc_ + 1, cix_ = nb_low(nb) ? 1 : nc_;      \
    int c2ix_ = nb_low(nb) ? 2 : nc_ - 1;                                 \
    int bmax_ = D_ == 3 ? nc_ : 1;                                        \
    for (int b_ = 1; b_ <= bmax_; b_++)                                   \
      for (int a_ = 1; a_ <= nc_; a_++) {                                 \
        int q_[3], r_[3], s_[3];                                          \
        int ta_ = (d_ == 0) ? 1 : 0, tb_ = (d_ == 2) ? 1 : 2;             \
        q_[d_] = gix_; r_[d_] = cix_; s_[d_] = c2ix_;                     \
        q_[ta_] = r_[ta_] = s_[ta_] = a_;                                 \
        if (D_ == 3) { q_[tb_] = r_[tb_] = s_[tb_] = b_; }                \
        size_t g = D_ == 2 ? I2(nc_, q_[0], q_[1]) : I3(nc_, q_[0], q_[1], q_[2]); \
        size_t c = D_ == 2 ? I2(nc_, r_[0], r_[1]) : I3(nc_, r_[0], r_[1], r_[2]); \
        size_t c2 = D_ == 2 ? I2(nc_, s_[0], s_[1]) : I3(nc_, s_[0], s_[1], s_[2]); \
        (void)c2;
#define FACE_LOOP_END \
  }                   \
  }

static void sides_from_nb(otree *t, int id, int nb_id, int nb, int iv) {
  /* sides_from_nb, m_aX_gc.f90:368-402: ghost = neighbour's adjacent layer */
  int nc = t->ncell[id];
  double *a = CC(t, id, iv), *o = CC(t, nb_id, iv);
  int d = nb_dim(nb);
  int src = nb_low(nb) ? nc : 1, dst = nb_low(nb) ? 0 : nc + 1;
  int bmax = t->dim == 3 ? nc : 1;
  for (int b = 1; b <= bmax; b++)
    for (int x = 1; x <= nc; x++) {
      int p[3], q[3];
      int ta = (d == 0) ? 1 : 0, tb = (d == 2) ? 1 : 2;
      p[d] = dst; q[d] = src;
      p[ta] = q[ta] = x;
      if (t->dim == 3) p[tb] = q[tb] = b;
      if (t->dim == 2) a[I2(nc, p[0], p[1])] = o[I2(nc, q[0], q[1])];
      else a[I3(nc, p[0], p[1], p[2])] = o[I3(nc, q[0], q[1], q[2])];
    }
}

static void bc_neumann(otree *t, int id, int nb, int iv) {
  /* a$D_gc_neumann, m_aX_gc.f90:298-330 */
  double *a = CC(t, id, iv);
  FACE_LOOP_BEGIN(t, id, nb)
  a[g] = a[c];
  FACE_LOOP_END
}

static void bc_dirichlet(otree *t, int id, int nb, int iv) {
  /* a$D_gc_dirichlet, m_aX_gc.f90:333-365 */
  double *a = CC(t, id, iv);
  FACE_LOOP_BEGIN(t, id, nb)
  a[g] = -a[c];
  FACE_LOOP_END
}

static void bc_mg(otree *t, int id, int nb, int iv) {
  /* The three forms of the drivers' potential boundary callbacks
   * (src/streamer_2d.f90:620-638, src/streamer_3d.f90:693-715,
   * src/streamer_cyl.f90:395-413): copy, -c, 2V - c on the face (kinds 0-2),
   * and 2V - c over the whole ghost row of a 2D y face (kind 3) */
  double *a = CC(t, id, iv);
  int kind = t->bc_kind[nb - 1];
  double v = t->bc_val[nb - 1];
  FACE_LOOP_BEGIN(t, id, nb)
  if (kind == 0) a[g] = a[c];
  else if (kind == 1) a[g] = -a[c];
  else a[g] = 2 * v - a[c];
  FACE_LOOP_END
  if (kind == 3 && t->dim == 2 && nb_dim(nb) == 1) {
    /* the cylindrical driver's rule covers the whole ghost row,
     * cc(:, nc+1) = 2V - cc(:, nc) (src/streamer_cyl.f90:408-410): the two
     * corner ghosts too, from the x ghosts of this pass */
    int nc = t->ncell[id], gy = nb_low(nb) ? 0 : nc + 1, cy = nb_low(nb) ? 1 : nc;
    a[I2(nc, 0, gy)] = 2 * v - a[I2(nc, 0, cy)];
    a[I2(nc, nc + 1, gy)] = 2 * v - a[I2(nc, nc + 1, cy)];
  }
}

static void mg_sides_rb(otree *t, int id, int nb, int iv) {
  /* mg$D_sides_rb, m_aX_mg.f90:262-356 */
  int nc = t->ncell[id];
  int ix = nb_low(nb) ? 1 : nc, dix = nb_low(nb) ? 1 : -1;
  {
    /* a$D_gc_prolong0, m_aX_gc.f90:71-86 */
    int lo[3] = {1, 1, 1}, hi[3] = {nc, nc, nc}, d = nb_dim(nb);
    lo[d] = nb_hi01(nb) * (nc + 1);
    hi[d] = nb_hi01(nb) * (nc + 1);
    prolong0_to(t, id, iv, lo, hi);
  }
  double *a = CC(t, id, iv);
  int d = nb_dim(nb);
  if (t->dim == 2) {
    if (d == 0) {
      int i = ix, di = dix;
      for (int j = 1; j <= nc; j++) {
        int dj = -1 + 2 * (j & 1);
        a[I2(nc, i - di, j)] = 0.5 * a[I2(nc, i - di, j)] + a[I2(nc, i, j)] -
                               0.25 * (a[I2(nc, i + di, j)] + a[I2(nc, i, j + dj)]);
      }
    } else {
      int j = ix, dj = dix;
      for (int i = 1; i <= nc; i++) {
        int di = -1 + 2 * (i & 1);
        a[I2(nc, i, j - dj)] = 0.5 * a[I2(nc, i, j - dj)] + a[I2(nc, i, j)] -
                               0.25 * (a[I2(nc, i + di, j)] + a[I2(nc, i, j + dj)]);
      }
    }
  } else {
    if (d == 0) {
      int i = ix, di = dix;
      for (int k = 1; k <= nc; k++) {
        int dk = -1 + 2 * (k & 1);
        for (int j = 1; j <= nc; j++) {
          int dj = -1 + 2 * (j & 1);
          a[I3(nc, i - di, j, k)] = 0.5 * a[I3(nc, i - di, j, k)] + 1.25 * a[I3(nc, i, j, k)] -
                                    0.25 * (a[I3(nc, i + di, j, k)] + a[I3(nc, i, j + dj, k)] +
                                            a[I3(nc, i, j, k + dk)]);
        }
      }
    } else if (d == 1) {
      int j = ix, dj = dix;
      for (int k = 1; k <= nc; k++) {
        int dk = -1 + 2 * (k & 1);
        for (int i = 1; i <= nc; i++) {
          int di = -1 + 2 * (i & 1);
          a[I3(nc, i, j - dj, k)] = 0.5 * a[I3(nc, i, j - dj, k)] + 1.25 * a[I3(nc, i, j, k)] -
                                    0.25 * (a[I3(nc, i + di, j, k)] + a[I3(nc, i, j + dj, k)] +
                                            a[I3(nc, i, j, k + dk)]);
        }
      }
    } else {
      int k = ix, dk = dix;
      for (int j = 1; j <= nc; j++) {
        int dj = -1 + 2 * (j & 1);
        for (int i = 1; i <= nc; i++) {
          int di = -1 + 2 * (i & 1);
          a[I3(nc, i, j, k - dk)] = 0.5 * a[I3(nc, i, j, k - dk)] + 1.25 * a[I3(nc, i, j, k)] -
                                    0.25 * (a[I3(nc, i + di, j, k)] + a[I3(nc, i, j + dj, k)] +
                                            a[I3(nc, i, j, k + dk)]);
        }
      }
    }
  }
}

static void gc_interp(otree *t, int id, int nb, int iv, int lim) {
  /* a$D_gc_interp (m_aX_gc.f90:90-185) and, with lim, a$D_gc_interp_lim
   * (m_aX_gc.f90:190-295) */
  const double sixth = 1 / 6.0, third = 1 / 3.0;
  int nc = t->ncell[id], p = t->parent[id], pn = NBP(t, p)[nb - 1], off[3];
  child_offset(t, id, nb, off);
  int ix, ixf, ixc;
  if (nb_low(nb)) { ix = 0; ixf = 1; ixc = nc; }
  else { ix = nc + 1; ixf = nc; ixc = 1; }
  double *a = CC(t, id, iv), *q = CC(t, pn, iv);
  int pnc = t->ncell[pn];
  int d = nb_dim(nb);
  if (t->dim == 2) {
    for (int x = 1; x <= nc; x++) {
      int o = (d == 0) ? off[1] : off[0];
      int c1 = o + ((x + 1) >> 1), c2 = c1 + 1 - 2 * (x & 1);
      double v1, v2;
      size_t g, f;
      if (d == 0) {
        v1 = q[I2(pnc, ixc, c1)]; v2 = q[I2(pnc, ixc, c2)];
        g = I2(nc, ix, x); f = I2(nc, ixf, x);
      } else {
        v1 = q[I2(pnc, c1, ixc)]; v2 = q[I2(pnc, c2, ixc)];
        g = I2(nc, x, ix); f = I2(nc, x, ixf);
      }
      a[g] = 0.5 * v1 + sixth * v2 + third * a[f];
      if (lim && a[g] > 2 * v1) a[g] = 2 * v1;
    }
  } else {
    for (int y = 1; y <= nc; y++)
      for (int x = 1; x <= nc; x++) {
        double v1, v2, v3;
        size_t g, f;
        if (d == 0) { /* loops k (y), j (x) */
          int k = y, j = x;
          int kc1 = off[2] + ((k + 1) >> 1), kc2 = kc1 + 1 - 2 * (k & 1);
          int jc1 = off[1] + ((j + 1) >> 1), jc2 = jc1 + 1 - 2 * (j & 1);
          v1 = q[I3(pnc, ixc, jc1, kc1)];
          v2 = q[I3(pnc, ixc, jc2, kc1)];
          v3 = q[I3(pnc, ixc, jc1, kc2)];
          g = I3(nc, ix, j, k); f = I3(nc, ixf, j, k);
        } else if (d == 1) { /* loops k (y), i (x) */
          int k = y, i = x;
          int kc1 = off[2] + ((k + 1) >> 1), kc2 = kc1 + 1 - 2 * (k & 1);
          int ic1 = off[0] + ((i + 1) >> 1), ic2 = ic1 + 1 - 2 * (i & 1);
          v1 = q[I3(pnc, ic1, ixc, kc1)];
          v2 = q[I3(pnc, ic2, ixc, kc1)];
          v3 = q[I3(pnc, ic1, ixc, kc2)];
          g = I3(nc, i, ix, k); f = I3(nc, i, ixf, k);
        } else { /* loops j (y), i (x) */
          int j = y, i = x;
          int jc1 = off[1] + ((j + 1) >> 1), jc2 = jc1 + 1 - 2 * (j & 1);
          int ic1 = off[0] + ((i + 1) >> 1), ic2 = ic1 + 1 - 2 * (i & 1);
          v1 = q[I3(pnc, ic1, jc1, ixc)];
          v2 = q[I3(pnc, ic1, jc2, ixc)];
          v3 = q[I3(pnc, ic2, jc1, ixc)];
          g = I3(nc, i, j, ix); f = I3(nc, i, j, ixf);
        }
        a[g] = third * v1 + sixth * v2 + sixth * v3 + third * a[f];
        if (lim && a[g] > 2 * v1) a[g] = 2 * v1;
      }
  }
}

void o_gc_box(otree *t, int id, int iv, int rb, int bc) {
  /* a$D_gc_box, m_aX_gc.f90:50-68 */
  for (int nb = 1; nb <= t->nnb; nb++) {
    int nb_id = NBP(t, id)[nb - 1];
    if (nb_id > NO_BOX) {
      sides_from_nb(t, id, nb_id, nb, iv);
    } else if (nb_id == NO_BOX) {
      if (rb == 0) mg_sides_rb(t, id, nb, iv);
      else gc_interp(t, id, nb, iv, rb == 2);
    } else {
      if (bc == 0) bc_mg(t, id, nb, iv);
      else if (bc == 1) bc_neumann(t, id, nb, iv);
      else bc_dirichlet(t, id, nb, iv);
    }
  }
}

void o_gc_tree(otree *t, int iv, int rb, int bc) {
  /* a$D_gc_tree, m_aX_gc.f90:28-46 */
  for (int lvl = t->min_lvl; lvl <= t->max_lvl; lvl++) {
    ivec *ids = &t->ids[LV(t, lvl)];
    for (int i = 0; i < ids->n; i++) o_gc_box(t, ids->v[i], iv, rb, bc);
  }
}

void o_gc2_box(otree *t, int id, int iv, double *out) {
  /* a$D_gc2_box with a$D_gc2_prolong1 / a$D_gc2_neumann,
   * m_aX_gc.f90:407-443, 446-482, 485-585, 588-622 */
  int nc = t->ncell[id], D = t->dim;
  size_t fl = D == 2 ? (size_t)nc : (size_t)nc * nc;
  double *a = CC(t, id, iv);
  for (int nb = 1; nb <= t->nnb; nb++) {
    double *o = out + (nb - 1) * fl;
    int nb_id = NBP(t, id)[nb - 1], d = nb_dim(nb);
    int ta = (d == 0) ? 1 : 0, tb = (d == 2) ? 1 : 2;
    if (nb_id > NO_BOX || nb_id < NO_BOX) {
      double *src = nb_id > NO_BOX ? CC(t, nb_id, iv) : a;
      int layer;
      if (nb_id > NO_BOX) layer = nb_low(nb) ? nc - 1 : 2;
      else layer = nb_low(nb) ? 2 : nc - 1;
      int bmax = D == 3 ? nc : 1;
      for (int b = 1; b <= bmax; b++)
        for (int x = 1; x <= nc; x++) {
          int p[3];
          p[d] = layer; p[ta] = x;
          if (D == 3) p[tb] = b;
          o[(size_t)(x - 1) + (size_t)nc * (b - 1)] =
              D == 2 ? src[I2(nc, p[0], p[1])] : src[I3(nc, p[0], p[1], p[2])];
        }
    } else {
      int p = t->parent[id], pn = NBP(t, p)[nb - 1], off[3], pnc;
      child_offset(t, id, nb, off);
      pnc = t->ncell[pn];
      double *q = CC(t, pn, iv);
      int ix = nb_hi01(nb) * (nc + 3) - 1;
      int n1 = off[d] + ((ix + 1) >> 1), n2 = n1 + 1 - 2 * (ix & 1);
      if (D == 2) {
        for (int x = 1; x <= nc; x++) {
          int c1 = off[ta] + ((x + 1) >> 1), c2 = c1 + 1 - 2 * (x & 1);
          double v1, v2, v3;
          if (d == 0) {
            v1 = q[I2(pnc, n1, c1)]; v2 = q[I2(pnc, n2, c1)]; v3 = q[I2(pnc, n1, c2)];
          } else {
            v1 = q[I2(pnc, c1, n1)]; v2 = q[I2(pnc, c2, n1)]; v3 = q[I2(pnc, c1, n2)];
          }
          o[x - 1] = 0.5 * v1 + 0.25 * v2 + 0.25 * v3;
        }
      } else {
        for (int b = 1; b <= nc; b++)
          for (int x = 1; x <= nc; x++) {
            int c1 = off[ta] + ((x + 1) >> 1), c2 = c1 + 1 - 2 * (x & 1);
            int e1 = off[tb] + ((b + 1) >> 1), e2 = e1 + 1 - 2 * (b & 1);
            int P[4][3];
            /* (n1,c1,e1), (n2,c1,e1), (n1,c2,e1), (n1,c1,e2) in the order of
             * the reference's i/j/k lists */
            for (int m = 0; m < 4; m++) { P[m][d] = n1; P[m][ta] = c1; P[m][tb] = e1; }
            P[1][d] = n2; P[2][ta] = c2; P[3][tb] = e2;
            /* the reference lists neighbours in (i,j,k) order: for d == 0 the
             * second term varies i, the third j, the fourth k; for d == 1:
             * i, j, k; for d == 2: i, j, k */
            double v[4];
            int ord[4];
            ord[0] = 0;
            /* term varying dimension 0, then 1, then 2 */
            int m = 1;
            for (int dd = 0; dd < 3; dd++) {
              if (dd == d) ord[m++] = 1;
              else if (dd == ta) ord[m++] = 2;
              else ord[m++] = 3;
            }
            for (int s = 0; s < 4; s++) v[s] = q[I3(pnc, P[ord[s]][0], P[ord[s]][1], P[ord[s]][2])];
            o[(size_t)(x - 1) + (size_t)nc * (b - 1)] = 0.25 * v[0] + 0.25 * v[1] + 0.25 * v[2] + 0.25 * v[3];
          }
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* restriction (m_aX_restrict.f90) and fluxes (m_aX_core.f90)           */
/* ------------------------------------------------------------------ */

static void restrict_box(otree *t, int c_id, int p_id, int iv, int i_dest) {
  /* a$D_restrict_box, m_aX_restrict.f90:67-142 */
  int nc = t->ncell[c_id], hnc = nc >> 1, pnc = t->ncell[p_id], off[3];
  child_offset(t, c_id, 0, off);
  double *c = CC(t, c_id, iv), *p = CC(t, p_id, i_dest);
  if (t->dim == 2) {
    if (t->coord == 1) {
      double dr16 = 0.0625 * t->dr[p_id];
      for (int j = 1; j <= hnc; j++) {
        int jc = off[1] + j, jf = 2 * j - 1;
        for (int i = 1; i <= hnc; i++) {
          int ic = off[0] + i, fi = 2 * i - 1;
          /* the radius is taken at parent cell (i, j), not (ic, jc), as in
           * the reference (m_aX_restrict.f90:106) */
          double r = RMP(t, p_id)[0] + (i - 0.5) * t->dr[p_id];
          double rfac = dr16 / r;
          double s1 = 0, s2 = 0;
          s1 = s1 + c[I2(nc, fi, jf)];
          s1 = s1 + c[I2(nc, fi, jf + 1)];
          s2 = s2 + c[I2(nc, fi + 1, jf)];
          s2 = s2 + c[I2(nc, fi + 1, jf + 1)];
          p[I2(pnc, ic, jc)] = (0.25 - rfac) * s1 + (0.25 + rfac) * s2;
        }
      }
    } else {
      for (int j = 1; j <= hnc; j++) {
        int jc = off[1] + j, jf = 2 * j - 1;
        for (int i = 1; i <= hnc; i++) {
          int ic = off[0] + i, fi = 2 * i - 1;
          double s = 0;
          s = s + c[I2(nc, fi, jf)];
          s = s + c[I2(nc, fi + 1, jf)];
          s = s + c[I2(nc, fi, jf + 1)];
          s = s + c[I2(nc, fi + 1, jf + 1)];
          p[I2(pnc, ic, jc)] = 0.25 * s;
        }
      }
    }
  } else {
    for (int k = 1; k <= hnc; k++) {
      int kc = off[2] + k, kf = 2 * k - 1;
      for (int j = 1; j <= hnc; j++) {
        int jc = off[1] + j, jf = 2 * j - 1;
        for (int i = 1; i <= hnc; i++) {
          int ic = off[0] + i, fi = 2 * i - 1;
          double s = 0;
          for (int dk = 0; dk < 2; dk++)
            for (int dj = 0; dj < 2; dj++)
              for (int di = 0; di < 2; di++) s = s + c[I3(nc, fi + di, jf + dj, kf + dk)];
          p[I3(pnc, ic, jc, kc)] = 0.125 * s;
        }
      }
    }
  }
}

void o_restrict_tree(otree *t, int iv) {
  /* a$D_restrict_tree, m_aX_restrict.f90:54-64 */
  for (int lvl = t->max_lvl - 1; lvl >= t->min_lvl; lvl--) {
    ivec *pa = &t->parents[LV(t, lvl)];
    for (int i = 0; i < pa->n; i++) {
      int id = pa->v[i];
      for (int c = 0; c < t->nch; c++) {
        int cid = CHP(t, id)[c];
        if (cid == NO_BOX) continue;
        restrict_box(t, cid, id, iv, iv);
      }
    }
  }
}

void o_consistent_fluxes(otree *t, int n, const int *f_ixs) {
  /* a$D_consistent_fluxes / flux_from_children, m_aX_core.f90:818-930 */
  for (int lvl = t->min_lvl; lvl <= t->max_lvl - 1; lvl++) {
    ivec *pa = &t->parents[LV(t, lvl)];
    for (int i = 0; i < pa->n; i++) {
      int id = pa->v[i];
      for (int nb = 1; nb <= t->nnb; nb++) {
        int nb_id = NBP(t, id)[nb - 1];
        if (nb_id <= NO_BOX || has_children(t, nb_id)) continue;
        int nc = t->ncell[id], nch = nc >> 1, d = nb_dim(nb);
        int ii = nb_low(nb) ? 1 : nc + 1, inb = nb_low(nb) ? nc + 1 : 1;
        for (int f = 0; f < n; f++) {
          int ifc = f_ixs[f];
          for (int ic = 0; ic < t->nch / 2; ic++) {
            int ich = ch_adj_nb(t, ic, nb), cid = CHP(t, id)[ich - 1];
            int ioff[3];
            for (int dd = 0; dd < t->dim; dd++) ioff[dd] = nch * ch_dix(ich, dd);
            double *dst = FC(t, nb_id, ifc, d), *src = FC(t, cid, ifc, d);
            int ta = (d == 0) ? 1 : 0, tb = (d == 2) ? 1 : 2;
            int bmax = t->dim == 3 ? nch : 1;
            for (int b = 1; b <= bmax; b++)
              for (int a = 1; a <= nch; a++) {
                int pd[3], s[3];
                pd[d] = inb; pd[ta] = ioff[ta] + a;
                s[d] = ii; s[ta] = 2 * a - 1;
                if (t->dim == 3) { pd[tb] = ioff[tb] + b; s[tb] = 2 * b - 1; }
                if (t->dim == 2) {
                  double v0 = src[FI(nc, 2, d, s[0], s[1], 1)];
                  s[ta] += 1;
                  double v1 = src[FI(nc, 2, d, s[0], s[1], 1)];
                  dst[FI(nc, 2, d, pd[0], pd[1], 1)] = 0.5 * (v0 + v1);
                } else {
                  int s0[3] = {s[0], s[1], s[2]};
                  double v0 = src[FI(nc, 3, d, s0[0], s0[1], s0[2])];
                  s0[ta] += 1;
                  double v1 = src[FI(nc, 3, d, s0[0], s0[1], s0[2])];
                  s0[ta] -= 1; s0[tb] += 1;
                  double v2 = src[FI(nc, 3, d, s0[0], s0[1], s0[2])];
                  s0[ta] += 1;
                  double v3 = src[FI(nc, 3, d, s0[0], s0[1], s0[2])];
                  dst[FI(nc, 3, d, pd[0], pd[1], pd[2])] = 0.25 * (v0 + v1 + v2 + v3);
                }
              }
          }
        }
      }
    }
  }
}

void o_tree_copy_cc(otree *t, int iv_from, int iv_to) {
  /* a$D_tree_copy_cc, m_aX_utils.f90:445-453 */
  for (int lvl = t->min_lvl; lvl <= t->max_lvl; lvl++) {
    ivec *ids = &t->ids[LV(t, lvl)];
    for (int i = 0; i < ids->n; i++) box_copy_cc(t, ids->v[i], iv_from, iv_to);
  }
}

void o_tree_minmax_cc(const otree *t, int iv, double *vmin, double *vmax) {
  /* a$D_tree_max_cc / a$D_tree_min_cc, m_aX_utils.f90:559-619 */
  double lo = 1.7976931348623157e308, hi = -1.7976931348623157e308;
  for (int lvl = t->min_lvl; lvl <= t->max_lvl; lvl++) {
    ivec *lv = &t->leaves[LV(t, lvl)];
    for (int i = 0; i < lv->n; i++) {
      int id = lv->v[i], nc = t->ncell[id];
      const double *a = CC(t, id, iv);
      int kmax = t->dim == 3 ? nc : 1;
      for (int k = 1; k <= kmax; k++)
        for (int j = 1; j <= nc; j++)
          for (int i2 = 1; i2 <= nc; i2++) {
            double v = t->dim == 2 ? a[I2(nc, i2, j)] : a[I3(nc, i2, j, k)];
            if (v < lo) lo = v;
            if (v > hi) hi = v;
          }
    }
  }
  *vmin = lo;
  *vmax = hi;
}

double o_min_dr(const otree *t) {
  /* a$D_min_dr / a$D_lvl_dr, m_aX_utils.f90:188-200 */
  return t->dr_base * pow(0.5, t->max_lvl - 1);
}

/* ------------------------------------------------------------------ */
/* locations and located reductions                                    */
/* ------------------------------------------------------------------ */

/* a$D_r_inside without the optional margin (m_aX_utils.f90:203-216) */
static int r_inside(const otree *t, int id, const double *r) {
  for (int d = 0; d < t->dim; d++) {
    double lo = RMP(t, id)[d], hi = lo + t->dr[id] * t->ncell[id];
    if (!(r[d] >= lo && r[d] <= hi)) return 0;
  }
  return 1;
}

void o_get_loc(const otree *t, const double *r, int max_lvl, int *id_out, int *ix) {
  /* a$D_get_loc (m_aX_utils.f90:244-281) with child_that_contains
     (:284-298) and a$D_cc_ix (:301-306); max_lvl < 1 stands for the absent
     argument (tree%lvls_max) */
  int lvl_max = max_lvl < 1 ? t->lvls_max : max_lvl;
  const ivec *l1 = &t->ids[LV(t, 1)];
  int id = -1;
  for (int i = 0; i < l1->n; i++)
    if (r_inside(t, l1->v[i], r)) {
      id = l1->v[i];
      break;
    }
  if (id < 0) {
    *id_out = -1;
    for (int d = 0; d < t->dim; d++) ix[d] = -1;
    return;
  }
  while (t->lvl[id] < lvl_max && CHP(t, id)[0] != 0) {
    int i_ch = 1;
    for (int d = 0; d < t->dim; d++) {
      double cntr = RMP(t, id)[d] + t->dr[id] * (t->ncell[id] >> 1);
      if (r[d] > cntr) i_ch += 1 << d;
    }
    id = CHP(t, id)[i_ch - 1];
  }
  *id_out = id;
  for (int d = 0; d < t->dim; d++) ix[d] = (int)ceil((r[d] - RMP(t, id)[d]) / t->dr[id]);
}

/* Per-box procedures of get_streamer_properties (src/streamer_cyl.f90):
   kind 1 is box_maxfld_r / box_maxfld_z (:831-842, :892-903) along
   dimension d, kind 2 is box_fld_maxr / box_fld_maxz (:844-890). */
static void box_rloc(const otree *t, int id, int kind, int var, int d, double thr,
                     double *val, int *ix) {
  int nc = t->ncell[id], kmax = t->dim == 3 ? nc : 1;
  for (int e = 0; e < t->dim; e++) ix[e] = -1;
  if (kind == 1) {
    /* maxloc(abs(f(1:nc along d) + f(2:nc+1 along d))): the first maximum in
       array order, value 0.5 * abs(...) there */
    const double *f = FC(t, id, var, d);
    double best = 0;
    int first = 1;
    for (int k = 1; k <= kmax; k++)
      for (int j = 1; j <= nc; j++)
        for (int i = 1; i <= nc; i++) {
          int q[3] = {i, j, k};
          q[d]++;
          double a = fabs(f[FI(nc, t->dim, d, i, j, k)] + f[FI(nc, t->dim, d, q[0], q[1], q[2])]);
          if (first || a > best) {
            best = a;
            ix[0] = i;
            ix[1] = j;
            if (t->dim == 3) ix[2] = k;
            first = 0;
          }
        }
    int q[3] = {ix[0], ix[1], t->dim == 3 ? ix[2] : 1};
    int p[3] = {q[0], q[1], q[2]};
    p[d]++;
    *val = 0.5 * fabs(f[FI(nc, t->dim, d, q[0], q[1], q[2])] + f[FI(nc, t->dim, d, p[0], p[1], p[2])]);
  } else {
    /* the largest cell-centre coordinate d (a$D_r_cc, m_aX_utils.f90:309-314)
       over cells with cc(var) > thr; 0 and no cell where none is */
    const double *c = CC(t, id, var);
    double v = 0;
    for (int k = 1; k <= kmax; k++)
      for (int j = 1; j <= nc; j++)
        for (int i = 1; i <= nc; i++) {
          size_t q = t->dim == 2 ? I2(nc, i, j) : I3(nc, i, j, k);
          if (c[q] > thr) {
            int w[3] = {i, j, k};
            double rc = RMP(t, id)[d] + (w[d] - 0.5) * t->dr[id];
            if (rc > v) {
              v = rc;
              ix[0] = i;
              ix[1] = j;
              if (t->dim == 3) ix[2] = k;
            }
          }
        }
    *val = v;
  }
}

void o_reduction_loc(const otree *t, int kind, int var, int d, double thr, double init_val,
                     double *val, int *id_out, int *ix) {
  /* a$D_reduction_loc (m_aX_utils.f90:499-555) with reduction = max
     (reduce_max, src/streamer_cyl.f90:826-829), in the order of one thread:
     levels upwards, leaves in list order; a box takes over where
     max(box value, running value) differs from the running value */
  double my = init_val;
  int my_id = -1, my_ix[3] = {-1, -1, -1};
  for (int lvl = t->min_lvl; lvl <= t->max_lvl; lvl++) {
    const ivec *lv = &t->leaves[LV(t, lvl)];
    for (int i = 0; i < lv->n; i++) {
      double v;
      int bix[3];
      box_rloc(t, lv->v[i], kind, var, d, thr, &v, bix);
      double nv = v > my ? v : my;
      if (fabs(nv - my) > 0) {
        my_id = lv->v[i];
        for (int e = 0; e < 3; e++) my_ix[e] = bix[e];
        my = v;
      }
    }
  }
  *val = my;
  *id_out = my_id;
  for (int e = 0; e < t->dim; e++) ix[e] = my_ix[e];
}

int o_cc_axis(const otree *t, int n_cc, const int *ixs_cc, int n_fc, const int *ixs_fc,
              double *out, int cap) {
  /* get_cc_axis (src/streamer_cyl.f90:905-956), 2D: the leaf boxes met
     walking up the line x = 0 from y = 0, each giving its first column of
     cells: y of the centre, cc(1, j, ixs_cc), 0.5 (fy(1, j) + fy(1, j+1))
     of ixs_fc. out(n_cc + n_fc + 1, n) column-major; returns n, writing
     only when out is given and n <= cap. */
  int nc = t->n_cell, rows = n_cc + n_fc + 1, cnt = 0;
  double r[2] = {0.0, 0.0};
  for (;;) {
    int id, ix[2];
    o_get_loc(t, r, 0, &id, ix);
    if (id == -1) break;
    double box_z = RMP(t, id)[1], box_dz = t->dr[id];
    if (out && cnt + nc <= cap) {
      for (int j = 1; j <= nc; j++) {
        double *o = out + (size_t)(cnt + j - 1) * rows;
        o[0] = box_z + (j - 0.5) * box_dz;
        for (int v = 0; v < n_cc; v++) o[1 + v] = CC(t, id, ixs_cc[v])[I2(nc, 1, j)];
        for (int v = 0; v < n_fc; v++) {
          const double *f = FC(t, id, ixs_fc[v], 1);
          o[1 + n_cc + v] = 0.5 * (f[FI(nc, 2, 1, 1, j, 0)] + f[FI(nc, 2, 1, 1, j + 1, 0)]);
        }
      }
    }
    cnt += nc;
    r[1] = box_z + (nc + 1) * box_dz;
  }
  return cnt;
}

long long o_leaf_cells(const otree *t, int n_cc, const int *ixs_cc, int n_fc,
                       const int *ixs_fc, double *out) {
  /* The cell data of a$D_write_vtk (m_aX_io.f90:87-174): leaves level by
     level in list order, cells x fastest; columns cc(ixs_cc(v)), then with
     face variables one column per dimension d: 0.5 * sum(f_d(two faces of
     the cell along d, ixs_fc)), the section summed in array element order
     from zero. out(n_cells, columns) column-major when given; returns
     n_cells. */
  int D = t->dim, nc = t->n_cell, cpb = D == 2 ? nc * nc : nc * nc * nc;
  long long n_cells = 0;
  for (int lvl = 1; lvl <= t->max_lvl; lvl++) n_cells += (long long)t->leaves[LV(t, lvl)].n * cpb;
  if (!out) return n_cells;
  long long c = 0;
  for (int lvl = 1; lvl <= t->max_lvl; lvl++) {
    const ivec *lv = &t->leaves[LV(t, lvl)];
    for (int n = 0; n < lv->n; n++) {
      int id = lv->v[n];
      for (int k = 1; k <= (D == 3 ? nc : 1); k++)
        for (int j = 1; j <= nc; j++)
          for (int i = 1; i <= nc; i++, c++) {
            size_t q = D == 2 ? I2(nc, i, j) : I3(nc, i, j, k);
            for (int v = 0; v < n_cc; v++) out[v * n_cells + c] = CC(t, id, ixs_cc[v])[q];
            if (!n_fc) continue;
            for (int d = 0; d < D; d++) {
              int lo[3] = {i, j, k}, hi[3] = {i, j, k};
              hi[d]++;
              double s = 0.0;
              for (int v = 0; v < n_fc; v++) {
                const double *f = FC(t, id, ixs_fc[v], d);
                s = s + f[FI(nc, D, d, lo[0], lo[1], lo[2])];
                s = s + f[FI(nc, D, d, hi[0], hi[1], hi[2])];
              }
              out[(n_cc + d) * n_cells + c] = 0.5 * s;
            }
          }
    }
  }
  return n_cells;
}

void o_vtk_geometry(const otree *t, double *coords, int *connects, int *offsets, int *types) {
  /* The node coordinates (dim per node), connectivity (0-based node
     indices, 2^dim per cell), offsets and cell types (8 pixel, 11 voxel) of
     a$D_write_vtk (m_aX_io.f90:106-174) */
  int D = t->dim, bc = t->n_cell, bn = bc + 1, n_ch = 1 << D;
  int npb = D == 2 ? bn * bn : bn * bn * bn, cpb = D == 2 ? bc * bc : bc * bc * bc;
  long long g = 0;
  for (int lvl = 1; lvl <= t->max_lvl; lvl++) {
    const ivec *lv = &t->leaves[LV(t, lvl)];
    for (int n = 0; n < lv->n; n++, g++) {
      int id = lv->v[n];
      long long node_ix = g * npb, cell_ix = g * cpb;
      for (int k = 1; k <= (D == 3 ? bn : 1); k++)
        for (int j = 1; j <= bn; j++)
          for (int i = 1; i <= bn; i++) {
            long long nn = node_ix + (long long)(k - 1) * bn * bn + (long long)(j - 1) * bn + i - 1;
            int w[3] = {i - 1, j - 1, k - 1};
            for (int d = 0; d < D; d++) coords[D * nn + d] = RMP(t, id)[d] + w[d] * t->dr[id];
          }
      for (int k = 1; k <= (D == 3 ? bc : 1); k++)
        for (int j = 1; j <= bc; j++)
          for (int i = 1; i <= bc; i++) {
            long long n_ix = node_ix + (long long)(k - 1) * bn * bn + (long long)(j - 1) * bn + i - 1;
            long long c_ix = cell_ix + (long long)(k - 1) * bc * bc + (long long)(j - 1) * bc + i;
            long long b2 = (long long)bn * bn;
            long long q[8] = {n_ix, n_ix + 1, n_ix + bn, n_ix + bn + 1,
                              n_ix + b2, n_ix + b2 + 1, n_ix + b2 + bn, n_ix + b2 + bn + 1};
            for (int e = 0; e < n_ch; e++) connects[n_ch * (c_ix - 1) + e] = (int)q[e];
            offsets[c_ix - 1] = (int)(n_ch * c_ix);
            types[c_ix - 1] = D == 2 ? 8 : 11;
          }
    }
  }
}

/* ------------------------------------------------------------------ */
/* multigrid (m_aX_mg.f90)                                              */
/* ------------------------------------------------------------------ */

void o_mg_setup(otree *t, int i_phi, int i_tmp, int i_rhs, int i_eps,
                int n_down, int n_up, int n_base, const int *bc_kind,
                const double *bc_val, int auto_ops) {
  /* mg$D_init_mg, m_aX_mg.f90:130-154 */
  t->i_phi = i_phi;
  t->i_tmp = i_tmp;
  t->i_rhs = i_rhs;
  t->i_eps = i_eps;
  t->n_down = n_down < 0 ? 2 : n_down;
  t->n_up = n_up < 0 ? 2 : n_up;
  t->n_base = n_base < 0 ? 2 : n_base;
  t->auto_ops = auto_ops;
  t->helm = 0;
  for (int nb = 0; nb < t->nnb; nb++) {
    t->bc_kind[nb] = bc_kind ? bc_kind[nb] : 0;
    t->bc_val[nb] = bc_val ? bc_val[nb] : 0;
  }
  t->mg_ok = 1;
}

void o_mg_set_bc_value(otree *t, int nb, double value) {
  /* the value V of the rule 2V - c on face nb (the drivers' sides_bc_pot
     reads ST_applied_voltage at each call, src/streamer_2d.f90:620-638) */
  t->bc_val[nb - 1] = value;
}

void o_mg_set_helmholtz(otree *t, double kappa2) {
  /* The Helmholtz operator lpl(phi) - kappa2 phi of the photoionization
   * model. No reference counterpart (the reference's photoionization is
   * Monte-Carlo); written from the discrete operator, the pattern of a custom
   * mg%box_op / mg%box_gsrb (afivo/examples/tutorial/implicit.f90:147-193).
   * Pinned by a direct sparse solve and by manufactured solutions. */
  t->helm = kappa2;
}

static int box_tag(otree *t, int id) {
  /* mg$D_set_box_tag, m_aX_mg.f90:673-712 (level-set boxes are not used by
   * the streamer drivers and are not restated) */
  if (!t->auto_ops) return TAG_NORMAL;
  if (t->tag[id] != INIT_TAG) return t->tag[id];
  int tag = TAG_NORMAL;
  if (t->i_eps != -1) {
    const double *e = CC(t, id, t->i_eps);
    int n = blen(t, id);
    double a = e[0], b = e[0];
    for (int i = 1; i < n; i++) {
      if (e[i] < a) a = e[i];
      if (e[i] > b) b = e[i];
    }
    if (b > a) tag = TAG_VEPS;
    else if (a < 1 || a > 1) tag = TAG_CEPS;
  }
  t->tag[id] = tag;
  return tag;
}

static void gsrb_lpl(otree *t, int id, int cntr, int cyl) {
  /* mg$D_box_gsrb_lpl (m_aX_mg.f90:769-811) and mg2_box_gsrb_clpl
   * (m_aX_mg.f90:1413-1439) */
  int nc = t->ncell[id];
  double dx2 = t->dr[id] * t->dr[id];
  double *p = CC(t, id, t->i_phi), *r = CC(t, id, t->i_rhs);
  if (t->dim == 2) {
    if (cyl) {
      int ioff = (IXP(t, id)[0] - 1) * nc;
      for (int j = 1; j <= nc; j++) {
        int i0 = 2 - ((cntr ^ j) & 1);
        for (int i = i0; i <= nc; i += 2) {
          double den = i + ioff - 0.5;
          double rf1 = (double)(i + ioff - 1) / den, rf2 = (double)(i + ioff) / den;
          double num = rf1 * p[I2(nc, i - 1, j)] + rf2 * p[I2(nc, i + 1, j)] +
                       p[I2(nc, i, j + 1)] + p[I2(nc, i, j - 1)] - dx2 * r[I2(nc, i, j)];
          p[I2(nc, i, j)] = t->helm != 0 ? num / (4 + t->helm * dx2) : 0.25 * num;
        }
      }
    } else {
      for (int j = 1; j <= nc; j++) {
        int i0 = 2 - ((cntr ^ j) & 1);
        for (int i = i0; i <= nc; i += 2) {
          double num = p[I2(nc, i + 1, j)] + p[I2(nc, i - 1, j)] + p[I2(nc, i, j + 1)] +
                       p[I2(nc, i, j - 1)] - dx2 * r[I2(nc, i, j)];
          p[I2(nc, i, j)] = t->helm != 0 ? num / (4 + t->helm * dx2) : 0.25 * num;
        }
      }
    }
  } else {
    const double sixth = 1 / 6.0;
    for (int k = 1; k <= nc; k++)
      for (int j = 1; j <= nc; j++) {
        int i0 = 2 - ((cntr ^ (k + j)) & 1);
        for (int i = i0; i <= nc; i += 2) {
          double num = p[I3(nc, i + 1, j, k)] + p[I3(nc, i - 1, j, k)] +
                       p[I3(nc, i, j + 1, k)] + p[I3(nc, i, j - 1, k)] +
                       p[I3(nc, i, j, k + 1)] + p[I3(nc, i, j, k - 1)] - dx2 * r[I3(nc, i, j, k)];
          p[I3(nc, i, j, k)] = t->helm != 0 ? num / (6 + t->helm * dx2) : sixth * num;
        }
      }
  }
}

static void gsrb_lpld(otree *t, int id, int cntr) {
  /* mg$D_box_gsrb_lpld (m_aX_mg.f90:895-948) and mg2_box_gsrb_clpld
   * (m_aX_mg.f90:1499-1531) */
  int nc = t->ncell[id], D = t->dim;
  double dx2 = t->dr[id] * t->dr[id];
  double *p = CC(t, id, t->i_phi), *r = CC(t, id, t->i_rhs), *e = CC(t, id, t->i_eps);
  int kmax = D == 3 ? nc : 1;
  int ioff = (IXP(t, id)[0] - 1) * nc;
  for (int k = 1; k <= kmax; k++)
    for (int j = 1; j <= nc; j++) {
      int i0 = 2 - ((cntr ^ (D == 3 ? k + j : j)) & 1);
      for (int i = i0; i <= nc; i += 2) {
        size_t c0 = D == 2 ? I2(nc, i, j) : I3(nc, i, j, k);
        size_t nbi[6];
        int nn = 2 * D;
        if (D == 2) {
          nbi[0] = I2(nc, i - 1, j); nbi[1] = I2(nc, i + 1, j);
          nbi[2] = I2(nc, i, j - 1); nbi[3] = I2(nc, i, j + 1);
        } else {
          nbi[0] = I3(nc, i - 1, j, k); nbi[1] = I3(nc, i + 1, j, k);
          nbi[2] = I3(nc, i, j - 1, k); nbi[3] = I3(nc, i, j + 1, k);
          nbi[4] = I3(nc, i, j, k - 1); nbi[5] = I3(nc, i, j, k + 1);
        }
        double a0 = e[c0], c[6], rf[6] = {1, 1, 1, 1, 1, 1};
        if (D == 2 && t->coord == 1) {
          double den = i + ioff - 0.5;
          rf[0] = (double)(i + ioff - 1) / den;
          rf[1] = (double)(i + ioff) / den;
        }
        for (int m = 0; m < nn; m++) c[m] = 2 * a0 * e[nbi[m]] / (a0 + e[nbi[m]]);
        double s1 = 0, s2 = 0;
        if (D == 2 && t->coord == 1) {
          for (int m = 0; m < nn; m++) s1 = s1 + c[m] * rf[m] * p[nbi[m]];
          for (int m = 0; m < nn; m++) s2 = s2 + c[m] * rf[m];
        } else {
          for (int m = 0; m < nn; m++) s1 = s1 + c[m] * p[nbi[m]];
          for (int m = 0; m < nn; m++) s2 = s2 + c[m];
        }
        p[c0] = (s1 - dx2 * r[c0]) / s2;
      }
    }
}

void o_box_gsrb(otree *t, int id, int cntr) {
  /* mg$D_auto_gsrb (m_aX_mg.f90:568-600) when the automatic operators are
   * installed; otherwise the library default mg$D_box_gsrb_lpl
   * (m_aX_mg.f90:149), which is Cartesian whatever the coordinates */
  int tag = box_tag(t, id);
  if (tag == TAG_NORMAL) gsrb_lpl(t, id, cntr, t->auto_ops && t->coord == 1);
  else gsrb_lpld(t, id, cntr);
}

static void box_lpl(otree *t, int id, int i_out, int cyl) {
  /* mg$D_box_lpl (m_aX_mg.f90:814-848), mg2_box_clpl (:1442-1464) */
  int nc = t->ncell[id];
  double inv_dr_sq = 1 / (t->dr[id] * t->dr[id]);
  double *p = CC(t, id, t->i_phi), *o = CC(t, id, i_out);
  if (t->dim == 2) {
    if (cyl) {
      int ioff = (IXP(t, id)[0] - 1) * nc;
      for (int j = 1; j <= nc; j++)
        for (int i = 1; i <= nc; i++) {
          double den = i + ioff - 0.5;
          double rf1 = (double)(i + ioff - 1) / den, rf2 = (double)(i + ioff) / den;
          o[I2(nc, i, j)] = (rf1 * p[I2(nc, i - 1, j)] + rf2 * p[I2(nc, i + 1, j)] +
                             p[I2(nc, i, j - 1)] + p[I2(nc, i, j + 1)] -
                             4 * p[I2(nc, i, j)]) * inv_dr_sq;
        }
    } else {
      for (int j = 1; j <= nc; j++)
        for (int i = 1; i <= nc; i++)
          o[I2(nc, i, j)] = inv_dr_sq * (p[I2(nc, i - 1, j)] + p[I2(nc, i + 1, j)] +
                                         p[I2(nc, i, j - 1)] + p[I2(nc, i, j + 1)] -
                                         4 * p[I2(nc, i, j)]);
    }
  } else {
    for (int k = 1; k <= nc; k++)
      for (int j = 1; j <= nc; j++)
        for (int i = 1; i <= nc; i++)
          o[I3(nc, i, j, k)] = inv_dr_sq * (p[I3(nc, i - 1, j, k)] + p[I3(nc, i + 1, j, k)] +
                                            p[I3(nc, i, j - 1, k)] + p[I3(nc, i, j + 1, k)] +
                                            p[I3(nc, i, j, k - 1)] + p[I3(nc, i, j, k + 1)] -
                                            6 * p[I3(nc, i, j, k)]);
  }
}

static void box_lpld(otree *t, int id, int i_out) {
  /* mg$D_box_lpld (m_aX_mg.f90:951-1001), mg2_box_clpld (:1467-1495) */
  int nc = t->ncell[id], D = t->dim;
  double inv_dr_sq = 1 / (t->dr[id] * t->dr[id]);
  double *p = CC(t, id, t->i_phi), *o = CC(t, id, i_out), *e = CC(t, id, t->i_eps);
  int kmax = D == 3 ? nc : 1;
  int ioff = (IXP(t, id)[0] - 1) * nc;
  for (int k = 1; k <= kmax; k++)
    for (int j = 1; j <= nc; j++)
      for (int i = 1; i <= nc; i++) {
        size_t c0 = D == 2 ? I2(nc, i, j) : I3(nc, i, j, k);
        size_t nbi[6];
        int nn = 2 * D;
        if (D == 2) {
          nbi[0] = I2(nc, i - 1, j); nbi[1] = I2(nc, i + 1, j);
          nbi[2] = I2(nc, i, j - 1); nbi[3] = I2(nc, i, j + 1);
        } else {
          nbi[0] = I3(nc, i - 1, j, k); nbi[1] = I3(nc, i + 1, j, k);
          nbi[2] = I3(nc, i, j - 1, k); nbi[3] = I3(nc, i, j + 1, k);
          nbi[4] = I3(nc, i, j, k - 1); nbi[5] = I3(nc, i, j, k + 1);
        }
        double u0 = p[c0], a0 = e[c0], s = 0;
        if (D == 2 && t->coord == 1) {
          double den = i + ioff - 0.5;
          double rf[4] = {(double)(i + ioff - 1) / den, (double)(i + ioff) / den, 1, 1};
          for (int m = 0; m < nn; m++)
            s = s + rf[m] * a0 * e[nbi[m]] / (a0 + e[nbi[m]]) * (p[nbi[m]] - u0);
        } else {
          for (int m = 0; m < nn; m++)
            s = s + a0 * e[nbi[m]] / (a0 + e[nbi[m]]) * (p[nbi[m]] - u0);
        }
        o[c0] = inv_dr_sq * 2 * s;
      }
}

void o_box_op(otree *t, int id, int i_out) {
  /* mg$D_auto_op (m_aX_mg.f90:603-635), or the default mg$D_box_lpl */
  if (box_tag(t, id) == TAG_NORMAL) {
    box_lpl(t, id, i_out, t->auto_ops && t->coord == 1);
    if (t->helm != 0) {
      /* the Helmholtz term: lpl(phi) - kappa2 phi */
      int nc = t->ncell[id], kmax = t->dim == 3 ? nc : 1;
      double *p = CC(t, id, t->i_phi), *o = CC(t, id, i_out);
      for (int k = 1; k <= kmax; k++)
        for (int j = 1; j <= nc; j++)
          for (int i = 1; i <= nc; i++) {
            size_t c = t->dim == 2 ? I2(nc, i, j) : I3(nc, i, j, k);
            o[c] = o[c] - t->helm * p[c];
          }
    }
  } else {
    box_lpld(t, id, i_out);
  }
}

void o_box_rstr(otree *t, int id, int iv) {
  /* mg$D_box_rstr_lpl, m_aX_mg.f90:851-893 (plain average, also in
   * cylindrical coordinates, as the default box_rstr) */
  int c_id = id, p_id = t->parent[id];
  int nc = t->ncell[c_id], hnc = nc >> 1, pnc = t->ncell[p_id], off[3];
  child_offset(t, c_id, 0, off);
  double *c = CC(t, c_id, iv), *p = CC(t, p_id, iv);
  if (t->dim == 2) {
    for (int j = 1; j <= hnc; j++) {
      int jc = off[1] + j, jf = 2 * j - 1;
      for (int i = 1; i <= hnc; i++) {
        int ic = off[0] + i, fi = 2 * i - 1;
        double s = 0;
        s = s + c[I2(nc, fi, jf)];
        s = s + c[I2(nc, fi + 1, jf)];
        s = s + c[I2(nc, fi, jf + 1)];
        s = s + c[I2(nc, fi + 1, jf + 1)];
        p[I2(pnc, ic, jc)] = 0.25 * s;
      }
    }
  } else {
    for (int k = 1; k <= hnc; k++) {
      int kc = off[2] + k, kf = 2 * k - 1;
      for (int j = 1; j <= hnc; j++) {
        int jc = off[1] + j, jf = 2 * j - 1;
        for (int i = 1; i <= hnc; i++) {
          int ic = off[0] + i, fi = 2 * i - 1;
          double s = 0;
          for (int dk = 0; dk < 2; dk++)
            for (int dj = 0; dj < 2; dj++)
              for (int di = 0; di < 2; di++) s = s + c[I3(nc, fi + di, jf + dj, kf + dk)];
          p[I3(pnc, ic, jc, kc)] = 0.125 * s;
        }
      }
    }
  }
}

void o_box_corr(otree *t, int id) {
  /* mg$D_auto_corr (m_aX_mg.f90:656-671) -> mg$D_box_corr_lpl (:714-766)
   * or mg$D_box_corr_lpld (:1003-1071) */
  int c_id = id, p_id = t->parent[id];
  int nc = t->ncell[c_id], pnc = t->ncell[p_id], off[3];
  int tag = box_tag(t, c_id);
  child_offset(t, c_id, 0, off);
  double *f = CC(t, c_id, t->i_phi), *q = CC(t, p_id, t->i_tmp);
  if (t->dim == 2) {
    const double *e = tag == TAG_NORMAL ? NULL : CC(t, p_id, t->i_eps);
    for (int j = 1; j <= nc; j++) {
      int jc1 = off[1] + ((j + 1) >> 1), jc2 = jc1 + 1 - 2 * (j & 1);
      for (int i = 1; i <= nc; i++) {
        int ic1 = off[0] + ((i + 1) >> 1), ic2 = ic1 + 1 - 2 * (i & 1);
        if (tag == TAG_NORMAL) {
          f[I2(nc, i, j)] = f[I2(nc, i, j)] + 0.5 * q[I2(pnc, ic1, jc1)] +
                            0.25 * (q[I2(pnc, ic2, jc1)] + q[I2(pnc, ic1, jc2)]);
        } else {
          double u0 = q[I2(pnc, ic1, jc1)], a0 = e[I2(pnc, ic1, jc1)];
          double u[2] = {q[I2(pnc, ic2, jc1)], q[I2(pnc, ic1, jc2)]};
          double a[2] = {e[I2(pnc, ic2, jc1)], e[I2(pnc, ic1, jc2)]};
          double s = 0;
          for (int m = 0; m < 2; m++) s = s + (a0 * u0 + a[m] * u[m]) / (a0 + a[m]);
          f[I2(nc, i, j)] = f[I2(nc, i, j)] + 0.5 * s;
        }
      }
    }
  } else {
    const double third = 1 / 3.0;
    const double *e = tag == TAG_NORMAL ? NULL : CC(t, p_id, t->i_eps);
    for (int k = 1; k <= nc; k++) {
      int kc1 = off[2] + ((k + 1) >> 1), kc2 = kc1 + 1 - 2 * (k & 1);
      for (int j = 1; j <= nc; j++) {
        int jc1 = off[1] + ((j + 1) >> 1), jc2 = jc1 + 1 - 2 * (j & 1);
        for (int i = 1; i <= nc; i++) {
          int ic1 = off[0] + ((i + 1) >> 1), ic2 = ic1 + 1 - 2 * (i & 1);
          if (tag == TAG_NORMAL) {
            f[I3(nc, i, j, k)] = f[I3(nc, i, j, k)] +
                                 0.25 * (q[I3(pnc, ic1, jc1, kc1)] + q[I3(pnc, ic2, jc1, kc1)] +
                                         q[I3(pnc, ic1, jc2, kc1)] + q[I3(pnc, ic1, jc1, kc2)]);
          } else {
            double u0 = q[I3(pnc, ic1, jc1, kc1)], a0 = e[I3(pnc, ic1, jc1, kc1)];
            double u[3] = {q[I3(pnc, ic2, jc1, kc1)], q[I3(pnc, ic1, jc2, kc1)],
                           q[I3(pnc, ic1, jc1, kc2)]};
            double a[3] = {e[I3(pnc, ic2, jc1, kc1)], e[I3(pnc, ic1, jc2, kc1)],
                           e[I3(pnc, ic1, jc1, kc2)]};
            double s = 0;
            for (int m = 0; m < 3; m++)
              s = s + (a0 * u0 + a[m] * (1.5 * u[m] - 0.5 * u0)) / (a0 + a[m]);
            f[I3(nc, i, j, k)] = f[I3(nc, i, j, k)] + third * s;
          }
        }
      }
    }
  }
}

static void residual_box(otree *t, int id) {
  /* residual_box, m_aX_mg.f90:551-565 */
  o_box_op(t, id, t->i_tmp);
  int nc = t->ncell[id];
  double *r = CC(t, id, t->i_rhs), *o = CC(t, id, t->i_tmp);
  int kmax = t->dim == 3 ? nc : 1;
  for (int k = 1; k <= kmax; k++)
    for (int j = 1; j <= nc; j++)
      for (int i = 1; i <= nc; i++) {
        size_t c = t->dim == 2 ? I2(nc, i, j) : I3(nc, i, j, k);
        o[c] = r[c] - o[c];
      }
}

static void fill_gc_phi(otree *t, int lvl) {
  /* fill_gc_phi, m_aX_mg.f90:358-371 */
  ivec *ids = &t->ids[LV(t, lvl)];
  for (int i = 0; i < ids->n; i++) o_gc_box(t, ids->v[i], t->i_phi, 0, 0);
}

static void correct_children(otree *t, int lvl) {
  /* correct_children, m_aX_mg.f90:374-401, over the parents of lvl */
  ivec *pa = &t->parents[LV(t, lvl)];
  for (int i = 0; i < pa->n; i++) {
    int id = pa->v[i];
    double *q = CC(t, id, t->i_tmp), *p = CC(t, id, t->i_phi);
    int n = blen(t, id);
    for (int m = 0; m < n; m++) q[m] = p[m] - q[m];
    for (int c = 0; c < t->nch; c++) {
      int cid = CHP(t, id)[c];
      if (cid == NO_BOX) continue;
      o_box_corr(t, cid);
    }
  }
}

static void gsrb_boxes(otree *t, int lvl, int n_cycle) {
  /* gsrb_boxes, m_aX_mg.f90:403-426 */
  ivec *ids = &t->ids[LV(t, lvl)];
  for (int n = 1; n <= 2 * n_cycle; n++) {
    for (int i = 0; i < ids->n; i++) o_box_gsrb(t, ids->v[i], n);
    for (int i = 0; i < ids->n; i++) o_gc_box(t, ids->v[i], t->i_phi, 0, 0);
  }
}

static void update_coarse(otree *t, int lvl) {
  /* update_coarse, m_aX_mg.f90:430-486 */
  ivec *ids = &t->ids[LV(t, lvl)];
  int nc = t->ncell[ids->v[0]];
  size_t nint = t->dim == 2 ? (size_t)nc * nc : (size_t)nc * nc * nc;
  double *save = malloc(sizeof(double) * nint);
  for (int i = 0; i < ids->n; i++) {
    int id = ids->v[i], p_id = t->parent[id];
    double *q = CC(t, id, t->i_tmp);
    int kmax = t->dim == 3 ? nc : 1;
    size_t m = 0;
    for (int k = 1; k <= kmax; k++)
      for (int j = 1; j <= nc; j++)
        for (int ii = 1; ii <= nc; ii++)
          save[m++] = q[t->dim == 2 ? I2(nc, ii, j) : I3(nc, ii, j, k)];
    residual_box(t, id);
    o_box_rstr(t, id, t->i_tmp);
    o_box_rstr(t, id, t->i_phi);
    (void)p_id;
    m = 0;
    for (int k = 1; k <= kmax; k++)
      for (int j = 1; j <= nc; j++)
        for (int ii = 1; ii <= nc; ii++)
          q[t->dim == 2 ? I2(nc, ii, j) : I3(nc, ii, j, k)] = save[m++];
  }
  free(save);
  fill_gc_phi(t, lvl - 1);
  ivec *pa = &t->parents[LV(t, lvl - 1)];
  for (int i = 0; i < pa->n; i++) {
    int id = pa->v[i];
    o_box_op(t, id, t->i_rhs);
    box_add_cc(t, id, t->i_tmp, t->i_rhs);
    box_copy_cc(t, id, t->i_phi, t->i_tmp);
  }
}

static void set_coarse_phi_rhs(otree *t, int lvl) {
  /* set_coarse_phi_rhs, m_aX_mg.f90:490-523 */
  if (lvl == t->max_lvl) fill_gc_phi(t, lvl);
  ivec *ids = &t->ids[LV(t, lvl)];
  for (int i = 0; i < ids->n; i++) {
    int id = ids->v[i];
    residual_box(t, id);
    o_box_rstr(t, id, t->i_tmp);
    o_box_rstr(t, id, t->i_phi);
  }
  fill_gc_phi(t, lvl - 1);
  ivec *pa = &t->parents[LV(t, lvl - 1)];
  for (int i = 0; i < pa->n; i++) {
    int id = pa->v[i];
    o_box_op(t, id, t->i_rhs);
    box_add_cc(t, id, t->i_tmp, t->i_rhs);
  }
}

static void init_phi_rhs(otree *t) {
  /* init_phi_rhs, m_aX_mg.f90:526-549: phi is cleared on all levels above
   * the coarsest, and rhs restricted down */
  for (int lvl = t->max_lvl; lvl >= t->min_lvl + 1; lvl--) {
    ivec *ids = &t->ids[LV(t, lvl)];
    for (int i = 0; i < ids->n; i++) {
      int id = ids->v[i];
      box_clear_cc(t, id, t->i_phi);
      o_box_rstr(t, id, t->i_rhs);
    }
  }
}

void o_fas_vcycle(otree *t, int max_lvl, int set_residual) {
  /* mg$D_fas_vcycle, m_aX_mg.f90:208-255 */
  int min_lvl = t->min_lvl;
  for (int lvl = max_lvl; lvl >= min_lvl + 1; lvl--) {
    gsrb_boxes(t, lvl, t->n_down);
    update_coarse(t, lvl);
  }
  gsrb_boxes(t, min_lvl, t->n_base);
  for (int lvl = min_lvl + 1; lvl <= max_lvl; lvl++) {
    correct_children(t, lvl - 1);
    fill_gc_phi(t, lvl);
    gsrb_boxes(t, lvl, t->n_up);
  }
  if (set_residual)
    for (int lvl = min_lvl; lvl <= max_lvl; lvl++) {
      ivec *ids = &t->ids[LV(t, lvl)];
      for (int i = 0; i < ids->n; i++) residual_box(t, ids->v[i]);
    }
}

void o_fas_fmg(otree *t, int set_residual, int first_call) {
  /* mg$D_fas_fmg, m_aX_mg.f90:164-203 */
  int min_lvl = t->min_lvl;
  if (first_call) init_phi_rhs(t);
  else
    for (int lvl = t->max_lvl; lvl >= min_lvl + 1; lvl--) set_coarse_phi_rhs(t, lvl);
  for (int lvl = min_lvl; lvl <= t->max_lvl; lvl++) {
    ivec *ids = &t->ids[LV(t, lvl)];
    for (int i = 0; i < ids->n; i++) box_copy_cc(t, ids->v[i], t->i_phi, t->i_tmp);
    if (lvl > min_lvl) {
      correct_children(t, lvl - 1);
      fill_gc_phi(t, lvl);
    }
    o_fas_vcycle(t, lvl, set_residual && lvl == t->max_lvl);
  }
}

double o_max_residual(otree *t) {
  double mx = 0;
  for (int lvl = t->min_lvl; lvl <= t->max_lvl; lvl++) {
    ivec *lv = &t->leaves[LV(t, lvl)];
    for (int i = 0; i < lv->n; i++) {
      int id = lv->v[i], nc = t->ncell[id];
      residual_box(t, id);
      double *o = CC(t, id, t->i_tmp);
      int kmax = t->dim == 3 ? nc : 1;
      for (int k = 1; k <= kmax; k++)
        for (int j = 1; j <= nc; j++)
          for (int ii = 1; ii <= nc; ii++) {
            double v = fabs(o[t->dim == 2 ? I2(nc, ii, j) : I3(nc, ii, j, k)]);
            if (v > mx) mx = v;
          }
    }
  }
  return mx;
}

double o_koren_mlim(double a, double b) {
  /* koren_mlim, src/m_streamer.f90:339-362 */
  const double sixth = 1 / 6.0;
  double aa = a * a, ab = a * b;
  if (ab <= 0) return 0;
  if (aa <= 0.25 * ab) return a;
  if (aa <= 2.5 * ab) return sixth * (b + 2 * a);
  return b;
}

/* ------------------------------------------------------------------ */
/* streamer physics                                                     */
/* ------------------------------------------------------------------ */

struct otable {
  int n_rows, n_cols;
  double x_min, inv_dx;
  double *rc; /* rows_cols(n_rows, n_cols), column-major */
};

otable *o_lt_create(double x_min, double x_max, int n_rows, int n_cols) {
  /* LT_create, fosito/m_lookup_table.f90:46-60 */
  otable *t = calloc(1, sizeof(otable));
  t->n_rows = n_rows;
  t->n_cols = n_cols;
  t->x_min = x_min;
  t->inv_dx = (n_rows - 1) / (x_max - x_min);
  t->rc = calloc((size_t)n_rows * n_cols, sizeof(double));
  return t;
}

void o_lt_free(otable *t) {
  if (!t) return;
  free(t->rc);
  free(t);
}

static int fi_adaptive(int n, const double *list, double val) {
  /* FI_adaptive_r (fosito/m_find_index.f90:56-67): smallest 1-based ix with
   * val <= list(ix), linear below 40 entries, bisection above */
  if (n < 40) {
    int ix;
    for (ix = 1; ix <= n; ix++)
      if (val <= list[ix - 1]) break;
    return ix;
  }
  int lo = 1, hi = n;
  while (lo < hi) {
    int mid = lo + (hi - lo) / 2;
    if (val <= list[mid - 1]) hi = mid;
    else lo = mid + 1;
  }
  if (val > list[lo - 1]) lo = n + 1;
  return lo;
}

static double lin_interp_list(int n, const double *x, const double *y, double xv) {
  /* LT_lin_interp_list, fosito/m_lookup_table.f90:201-221 */
  if (xv <= x[0]) return y[0];
  if (xv >= x[n - 1]) return y[n - 1];
  int ix = fi_adaptive(n, x, xv);
  double temp = (xv - x[ix - 2]) / (x[ix - 1] - x[ix - 2]);
  return (1 - temp) * y[ix - 2] + temp * y[ix - 1];
}

void o_lt_set_col(otable *t, int col, int n, const double *xx, const double *yy) {
  /* LT_set_col / get_spaced_data / LT_get_xdata,
   * fosito/m_lookup_table.f90:62-72, 180-198 */
  double dx = 1 / t->inv_dx;
  for (int ix = 1; ix <= t->n_rows; ix++) {
    double xv = t->x_min + (ix - 1) * dx;
    t->rc[(size_t)(col - 1) * t->n_rows + (ix - 1)] = lin_interp_list(n, xx, yy, xv);
  }
}

static void lt_loc(const otable *t, double x, int *low_ix, double *low_frac) {
  /* get_loc, fosito/m_lookup_table.f90:107-125 */
  double frac = (x - t->x_min) * t->inv_dx;
  int li = (int)ceil(frac);
  double lf = li - frac;
  if (li < 1) {
    li = 1;
    lf = 1;
  } else if (li >= t->n_rows) {
    li = t->n_rows - 1;
    lf = 0;
  }
  *low_ix = li;
  *low_frac = lf;
}

static double lt_col_at(const otable *t, int col, int li, double lf) {
  /* LT_get_col_at_loc, fosito/m_lookup_table.f90:154-161 */
  const double *c = t->rc + (size_t)(col - 1) * t->n_rows;
  return lf * c[li - 1] + (1 - lf) * c[li];
}

double o_lt_get_col(const otable *t, int col, double x) {
  int li;
  double lf;
  lt_loc(t, x, &li, &lf);
  return lt_col_at(t, col, li, lf);
}

void o_lt_data(const otable *t, double *rc, double *x_min, double *inv_dx) {
  memcpy(rc, t->rc, sizeof(double) * (size_t)t->n_rows * t->n_cols);
  *x_min = t->x_min;
  *inv_dx = t->inv_dx;
}

double o_norm2(int n, const double *x) {
  /* NORM2 as the flang runtime evaluates it: running maximum m and
   * s = sum((x/m)^2) over the others, result m sqrt(1 + s) */
  double mx = 0, s = 0;
  for (int i = 0; i < n; i++) {
    double ax = fabs(x[i]);
    if (mx == 0) {
      mx = ax;
    } else if (ax > mx) {
      double q = mx / ax, qq = q * q;
      s = s * qq;
      s = s + qq;
      mx = ax;
    } else {
      double q = ax / mx;
      s = s + q * q;
    }
  }
  return mx * sqrt(1 + s);
}

void o_space_charge(otree *t, const o_vars *v, double fac) {
  /* compute_fld (src/streamer_2d.f90:318-329): leaves of levels >= 1 */
  for (int lvl = 1; lvl <= t->max_lvl; lvl++) {
    ivec *lv = &t->leaves[LV(t, lvl)];
    for (int i = 0; i < lv->n; i++) {
      int id = lv->v[i], n = blen(t, id);
      double *r = CC(t, id, v->rhs), *e = CC(t, id, v->elec), *p = CC(t, id, v->pion);
      for (int m = 0; m < n; m++) r[m] = fac * (e[m] - p[m]);
    }
  }
}

static void fld_box(otree *t, const o_vars *v, int id) {
  /* fld_from_pot, src/streamer_2d.f90:346-381, src/streamer_3d.f90:348-395,
   * src/streamer_cyl.f90:356-392 */
  int nc = t->ncell[id], D = t->dim;
  double inv_dr = 1 / t->dr[id];
  const double *p = CC(t, id, v->phi), *e = CC(t, id, v->eps);
  double *fl = CC(t, id, v->fld);
  double *F[3];
  for (int d = 0; d < D; d++) F[d] = FC(t, id, v->f_fld, d);
  int kmax = D == 3 ? nc : 1;
  for (int d = 0; d < D; d++) {
    int n0 = nc + (d == 0), n1 = nc + (d == 1), n2 = D == 3 ? nc + (d == 2) : 1;
    for (int k = 1; k <= n2; k++)
      for (int j = 1; j <= n1; j++)
        for (int i = 1; i <= n0; i++) {
          int a[3] = {i, j, D == 3 ? k : 0}, b[3] = {i, j, D == 3 ? k : 0};
          a[d] -= 1; /* cell on the low side of face (i, j, k) */
          size_t ia = D == 2 ? I2(nc, a[0], a[1]) : I3(nc, a[0], a[1], a[2]);
          size_t ib = D == 2 ? I2(nc, b[0], b[1]) : I3(nc, b[0], b[1], b[2]);
          int fpos = (d == 0) ? i : (d == 1) ? j : k;
          double val;
          if (fpos == 1) {
            val = 2 * inv_dr * (p[ia] - p[ib]) * e[ia] / (e[ib] + e[ia]);
          } else if (fpos == nc + 1) {
            val = 2 * inv_dr * (p[ia] - p[ib]) * e[ib] / (e[ib] + e[ia]);
          } else {
            val = inv_dr * (p[ia] - p[ib]);
          }
          F[d][FI(nc, D, d, i, j, k)] = val;
        }
  }
  for (int k = 1; k <= kmax; k++)
    for (int j = 1; j <= nc; j++)
      for (int i = 1; i <= nc; i++) {
        double sx = F[0][FI(nc, D, 0, i, j, k)] + F[0][FI(nc, D, 0, i + 1, j, k)];
        double sy = F[1][FI(nc, D, 1, i, j, k)] + F[1][FI(nc, D, 1, i, j + 1, k)];
        if (D == 2) {
          fl[I2(nc, i, j)] = sqrt(0.25 * (sx * sx) + 0.25 * (sy * sy));
        } else {
          double sz = F[2][FI(nc, D, 2, i, j, k)] + F[2][FI(nc, D, 2, i, j, k + 1)];
          fl[I3(nc, i, j, k)] = 0.5 * sqrt(sx * sx + sy * sy + sz * sz);
        }
      }
}

void o_fld_from_pot(otree *t, const o_vars *v) {
  /* a$D_loop_box(tree, fld_from_pot): every box of every level */
  for (int lvl = t->min_lvl; lvl <= t->max_lvl; lvl++) {
    ivec *ids = &t->ids[LV(t, lvl)];
    for (int i = 0; i < ids->n; i++) fld_box(t, v, ids->v[i]);
  }
}

static double koren_face(double vd, double c_up, double gradc, double g_other, double fac,
                         double fld, int neg) {
  /* one advective face flux of fluxes_koren (src/streamer_2d.f90:415-437) */
  double f;
  if (neg) {
    f = vd * (c_up - o_koren_mlim(gradc, g_other));
    if (f < fac * fld) f = fac * fld;
  } else {
    f = vd * (c_up + o_koren_mlim(gradc, g_other));
    if (f > fac * fld) f = fac * fld;
  }
  return f;
}

void o_fluxes_koren(otree *t, const o_vars *v, const otable *tbl, double dt) {
  /* fluxes_koren, src/streamer_2d.f90:384-489 (3D: src/streamer_3d.f90:396-554,
   * cylindrical: src/streamer_cyl.f90:416-521), on leaves */
  const double eps0 = 8.8541878176e-12, qe = 1.6022e-19;
  double fac = -0.8 * eps0 / (qe * dt);
  int D = t->dim;
  for (int lvl = t->min_lvl; lvl <= t->max_lvl; lvl++) {
    ivec *lv = &t->leaves[LV(t, lvl)];
    for (int q = 0; q < lv->n; q++) {
      int id = lv->v[q], nc = t->ncell[id];
      double inv_dr = 1 / t->dr[id];
      size_t fl = D == 2 ? (size_t)nc : (size_t)nc * nc;
      double *gc = malloc(sizeof(double) * fl * t->nnb);
      o_gc2_box(t, id, v->elec, gc);
      const double *n = CC(t, id, v->elec), *E = CC(t, id, v->fld);
      int kmax = D == 3 ? nc : 1;
      for (int d = 0; d < D; d++) {
        double *Fe = FC(t, id, v->f_elec, d), *Ff = FC(t, id, v->f_fld, d);
        int n0 = nc + (d == 0), n1 = nc + (d == 1), n2 = D == 3 ? nc + (d == 2) : 1;
        int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
        for (int k = 1; k <= n2; k++)
          for (int j = 1; j <= n1; j++)
            for (int i = 1; i <= n0; i++) {
              int c[3] = {i, j, D == 3 ? k : 0};
#define CI(a0, a1, a2) (D == 2 ? I2(nc, a0, a1) : I3(nc, a0, a1, a2))
              int m1[3] = {c[0], c[1], c[2]}, m2[3] = {c[0], c[1], c[2]}, p1[3] = {c[0], c[1], c[2]};
              m1[d] -= 1;
              m2[d] -= 2;
              p1[d] += 1;
              int pos = c[d];
              double fld_avg = 0.5 * (E[CI(c[0], c[1], c[2])] + E[CI(m1[0], m1[1], m1[2])]);
              int li;
              double lf;
              lt_loc(tbl, fld_avg, &li, &lf);
              double mob = lt_col_at(tbl, v->t_mob, li, lf);
              double dif = lt_col_at(tbl, v->t_dif, li, lf);
              size_t fi = FI(nc, D, d, i, j, k);
              double fld = Ff[fi];
              double vd = -mob * fld;
              double nc0 = n[CI(c[0], c[1], c[2])], nm1 = n[CI(m1[0], m1[1], m1[2])];
              double gradc = nc0 - nm1;
              size_t gpos = (size_t)(c[ta] - 1) + (D == 3 ? (size_t)nc * (c[tb] - 1) : 0);
              double f;
              if (vd < 0.0) {
                double tmp = pos == nc + 1 ? gc[(size_t)(2 * d + 1) * fl + gpos]
                                           : n[CI(p1[0], p1[1], p1[2])];
                f = koren_face(vd, nc0, gradc, tmp - nc0, fac, fld, 1);
              } else {
                double tmp = pos == 1 ? gc[(size_t)(2 * d) * fl + gpos]
                                      : n[CI(m2[0], m2[1], m2[2])];
                f = koren_face(vd, nm1, gradc, nm1 - tmp, fac, fld, 0);
              }
              Fe[fi] = f - dif * gradc * inv_dr;
#undef CI
            }
      }
      (void)kmax;
      free(gc);
    }
  }
}

void o_photoi_rate(otree *t, const o_vars *v, const otable *tbl, double coeff) {
  /* set_photoi_rate, src/streamer_2d.f90:567-589: |E| mu alpha n_e coeff on
   * leaves, clamped at zero (3D and cylindrical drivers alike) */
  for (int lvl = t->min_lvl; lvl <= t->max_lvl; lvl++) {
    ivec *lv = &t->leaves[LV(t, lvl)];
    for (int q = 0; q < lv->n; q++) {
      int id = lv->v[q], nc = t->ncell[id], kmax = t->dim == 3 ? nc : 1;
      const double *E = CC(t, id, v->fld), *ne = CC(t, id, v->elec);
      double *out = CC(t, id, v->pho);
      for (int k = 1; k <= kmax; k++)
        for (int j = 1; j <= nc; j++)
          for (int i = 1; i <= nc; i++) {
            size_t c = t->dim == 2 ? I2(nc, i, j) : I3(nc, i, j, k);
            int li;
            double lf;
            lt_loc(tbl, E[c], &li, &lf);
            double alpha = lt_col_at(tbl, v->t_alpha, li, lf);
            double mob = lt_col_at(tbl, v->t_mob, li, lf);
            double tmp = E[c] * mob * alpha * ne[c] * coeff;
            if (tmp < 0) tmp = 0;
            out[c] = tmp;
          }
    }
  }
}

void o_axpy_leaves(otree *t, int iin, int iout, double s, int add) {
  /* out = s in (add = 0) or out += s in on leaf interiors */
  for (int lvl = t->min_lvl; lvl <= t->max_lvl; lvl++) {
    ivec *lv = &t->leaves[LV(t, lvl)];
    for (int q = 0; q < lv->n; q++) {
      int id = lv->v[q], nc = t->ncell[id], kmax = t->dim == 3 ? nc : 1;
      const double *x = CC(t, id, iin);
      double *y = CC(t, id, iout);
      for (int k = 1; k <= kmax; k++)
        for (int j = 1; j <= nc; j++)
          for (int i = 1; i <= nc; i++) {
            size_t c = t->dim == 2 ? I2(nc, i, j) : I3(nc, i, j, k);
            y[c] = add ? y[c] + s * x[c] : s * x[c];
          }
    }
  }
}

static double leaf_maxabs(otree *t, int iv) {
  double mx = 0;
  for (int lvl = t->min_lvl; lvl <= t->max_lvl; lvl++) {
    ivec *lv = &t->leaves[LV(t, lvl)];
    for (int q = 0; q < lv->n; q++) {
      int id = lv->v[q], nc = t->ncell[id], kmax = t->dim == 3 ? nc : 1;
      const double *x = CC(t, id, iv);
      for (int k = 1; k <= kmax; k++)
        for (int j = 1; j <= nc; j++)
          for (int i = 1; i <= nc; i++) {
            double a = fabs(x[t->dim == 2 ? I2(nc, i, j) : I3(nc, i, j, k)]);
            if (a > mx) mx = a;
          }
    }
  }
  return mx;
}

int o_photoi_helmholtz(otree *t, const o_vars *v, const otable *tbl, double coeff,
                       int n_terms, const double *amp, const double *kappa, int i_phi,
                       int i_rhs, int i_tmp, double tol, int max_cycles, int *cycles) {
  /* The three-term Helmholtz photoionization model (Bourdon et al. 2007) in
   * place of the reference's Monte-Carlo photons: no reference counterpart,
   * restated from the model (see o_mg_set_helmholtz). Returns 0, or -1 when a
   * solve does not converge within max_cycles V-cycles. */
  o_photoi_rate(t, v, tbl, coeff);
  int s_phi = t->i_phi, s_rhs = t->i_rhs, s_tmp = t->i_tmp, s_eps = t->i_eps, s_ok = t->mg_ok;
  int s_kind[6];
  double s_val[6], s_helm = t->helm;
  for (int i = 0; i < 6; i++) {
    s_kind[i] = t->bc_kind[i];
    s_val[i] = t->bc_val[i];
    t->bc_kind[i] = 1;
    t->bc_val[i] = 0;
  }
  t->i_phi = i_phi;
  t->i_rhs = i_rhs;
  t->i_tmp = i_tmp;
  t->i_eps = -1;
  t->mg_ok = 1;
  int total = 0, status = 0;
  o_axpy_leaves(t, v->pho, i_rhs, -amp[0], 0);
  o_axpy_leaves(t, v->pho, v->pho, 0.0, 0);
  for (int j = 0; j < n_terms && status == 0; j++) {
    if (j > 0) o_axpy_leaves(t, i_rhs, i_rhs, amp[j] / amp[j - 1], 0);
    t->helm = kappa[j] * kappa[j];
    o_fas_fmg(t, 0, 1);
    for (int cyc = 0;; cyc++) {
      double res = o_max_residual(t), rhs = leaf_maxabs(t, i_rhs);
      if (res <= tol * rhs) break;
      if (cyc >= max_cycles) {
        status = -1;
        break;
      }
      o_fas_vcycle(t, t->max_lvl, 0);
      total++;
    }
    if (status == 0) o_axpy_leaves(t, i_phi, v->pho, 1.0, 1);
  }
  t->i_phi = s_phi;
  t->i_rhs = s_rhs;
  t->i_tmp = s_tmp;
  t->i_eps = s_eps;
  t->mg_ok = s_ok;
  t->helm = s_helm;
  for (int i = 0; i < 6; i++) {
    t->bc_kind[i] = s_kind[i];
    t->bc_val[i] = s_val[i];
  }
  if (cycles) *cycles = total;
  return status;
}

int o_photoi_helmholtz_terms(otree *t, const o_vars *v, const otable *tbl, double coeff,
                             int n_terms, const double *amp, const double *kappa,
                             const int *i_terms, const int *i_rhs, const int *i_tmp, int warm,
                             double tol, int max_cycles, int *cycles) {
  /* The same model with term j in variables of its own: S_j in i_terms[j],
   * rhs_j = -amp_j I in i_rhs[j], scratch in i_tmp[j]; the terms are solved
   * one after the other here (they share no data). Cold: FMG from zero
   * (first_call). Warm: i_terms[j] holds an earlier solution; it is
   * restricted to every parent (o_box_rstr, the multigrid's restriction),
   * ghost cells are filled level by level upwards, and V-cycles alone reach
   * the tolerance. max|rhs_j| is taken once: a FAS solve leaves the leaf rhs
   * as it is. Returns 0, or -1 when a solve does not converge. */
  o_photoi_rate(t, v, tbl, coeff);
  int s_phi = t->i_phi, s_rhs = t->i_rhs, s_tmp = t->i_tmp, s_eps = t->i_eps, s_ok = t->mg_ok;
  int s_kind[6];
  double s_val[6], s_helm = t->helm;
  for (int i = 0; i < 6; i++) {
    s_kind[i] = t->bc_kind[i];
    s_val[i] = t->bc_val[i];
    t->bc_kind[i] = 1;
    t->bc_val[i] = 0;
  }
  t->i_eps = -1;
  t->mg_ok = 1;
  int total = 0, status = 0;
  for (int j = 0; j < n_terms; j++) o_axpy_leaves(t, v->pho, i_rhs[j], -amp[j], 0);
  o_axpy_leaves(t, v->pho, v->pho, 0.0, 0);
  for (int j = 0; j < n_terms && status == 0; j++) {
    t->i_phi = i_terms[j];
    t->i_rhs = i_rhs[j];
    t->i_tmp = i_tmp[j];
    t->helm = kappa[j] * kappa[j];
    if (warm) {
      for (int lvl = t->max_lvl; lvl >= t->min_lvl + 1; lvl--) {
        ivec *ids = &t->ids[LV(t, lvl)];
        for (int i = 0; i < ids->n; i++) o_box_rstr(t, ids->v[i], t->i_phi);
      }
      for (int lvl = t->min_lvl; lvl <= t->max_lvl; lvl++) fill_gc_phi(t, lvl);
    } else {
      o_fas_fmg(t, 0, 1);
    }
    double rhs = leaf_maxabs(t, i_rhs[j]);
    for (int cyc = 0;; cyc++) {
      double res = o_max_residual(t);
      if (res <= tol * rhs) break;
      if (cyc >= max_cycles) {
        status = -1;
        break;
      }
      o_fas_vcycle(t, t->max_lvl, 0);
      total++;
    }
  }
  if (status == 0)
    for (int j = 0; j < n_terms; j++) o_axpy_leaves(t, i_terms[j], v->pho, 1.0, 1);
  t->i_phi = s_phi;
  t->i_rhs = s_rhs;
  t->i_tmp = s_tmp;
  t->i_eps = s_eps;
  t->mg_ok = s_ok;
  t->helm = s_helm;
  for (int i = 0; i < 6; i++) {
    t->bc_kind[i] = s_kind[i];
    t->bc_val[i] = s_val[i];
  }
  if (cycles) *cycles = total;
  return status;
}

void o_update_solution(otree *t, const o_vars *v, const otable *tbl, double dt) {
  /* update_solution, src/streamer_2d.f90:499-542 (3D :566-615, cylindrical
   * src/streamer_cyl.f90:531-570), on leaves */
  int D = t->dim;
  for (int lvl = t->min_lvl; lvl <= t->max_lvl; lvl++) {
    ivec *lv = &t->leaves[LV(t, lvl)];
    for (int q = 0; q < lv->n; q++) {
      int id = lv->v[q], nc = t->ncell[id];
      double inv_dr = 1 / t->dr[id];
      double *ne = CC(t, id, v->elec), *ni = CC(t, id, v->pion);
      const double *E = CC(t, id, v->fld), *eps = CC(t, id, v->eps);
      const double *ph = v->photoi ? CC(t, id, v->pho) : NULL;
      const double *F[3];
      for (int d = 0; d < D; d++) F[d] = FC(t, id, v->f_elec, d);
      int ioff = (IXP(t, id)[0] - 1) * nc;
      int kmax = D == 3 ? nc : 1;
      for (int k = 1; k <= kmax; k++)
        for (int j = 1; j <= nc; j++)
          for (int i = 1; i <= nc; i++) {
            size_t c = D == 2 ? I2(nc, i, j) : I3(nc, i, j, k);
            int li;
            double lf;
            lt_loc(tbl, E[c], &li, &lf);
            double alpha = lt_col_at(tbl, v->t_alpha, li, lf);
            double eta = lt_col_at(tbl, v->t_eta, li, lf);
            double fxl = F[0][FI(nc, D, 0, i, j, k)], fxh = F[0][FI(nc, D, 0, i + 1, j, k)];
            double fyl = F[1][FI(nc, D, 1, i, j, k)], fyh = F[1][FI(nc, D, 1, i, j + 1, k)];
            double df[3];
            df[0] = fxl + fxh;
            df[1] = fyl + fyh;
            double fzl = 0, fzh = 0;
            if (D == 3) {
              fzl = F[2][FI(nc, D, 2, i, j, k)];
              fzh = F[2][FI(nc, D, 2, i, j, k + 1)];
              df[2] = fzl + fzh;
            }
            double src = 0.5 * o_norm2(D, df) * (alpha - eta);
            if (ph) src = src + ph[c];
            double sflux;
            if (D == 2 && t->coord == 1) {
              double den = i + ioff - 0.5;
              double rf1 = (double)(i + ioff - 1) / den, rf2 = (double)(i + ioff) / den;
              sflux = (fyl - fyh + rf1 * fxl - rf2 * fxh) * inv_dr;
            } else if (D == 2) {
              sflux = (fxl - fxh + fyl - fyh) * inv_dr;
            } else {
              sflux = (fxl - fxh + fyl - fyh + fzl - fzh) * inv_dr;
            }
            ne[c] = ne[c] + (src + sflux) * dt;
            ni[c] = ni[c] + src * dt;
          }
      if (v->dielectric_rule) {
        for (int k = 1; k <= kmax; k++)
          for (int j = 1; j <= nc; j++)
            for (int i = 1; i <= nc; i++) {
              size_t c = D == 2 ? I2(nc, i, j) : I3(nc, i, j, k);
              if (eps[c] > 1.0) {
                ni[c] = ni[c] - ne[c];
                ne[c] = 0;
              }
            }
      }
    }
  }
}

void o_average_dens(otree *t, const o_vars *v) {
  /* average_dens (src/streamer_2d.f90:492-496) on every box, whole boxes */
  for (int lvl = t->min_lvl; lvl <= t->max_lvl; lvl++) {
    ivec *ids = &t->ids[LV(t, lvl)];
    for (int q = 0; q < ids->n; q++) {
      int id = ids->v[q], n = blen(t, id);
      double *e = CC(t, id, v->elec), *eo = CC(t, id, v->elec_old);
      double *p = CC(t, id, v->pion), *po = CC(t, id, v->pion_old);
      for (int m = 0; m < n; m++) e[m] = 0.5 * (e[m] + eo[m]);
      for (int m = 0; m < n; m++) p[m] = 0.5 * (p[m] + po[m]);
    }
  }
}

void o_box_summaries(const otree *t, const o_vars *v, double *out) {
  /* the maxima set_ref_flags takes per box (src/streamer_2d.f90:186-192) */
  for (int id = 1; id <= t->max_id; id++) {
    out[(id - 1) * 3 + 0] = out[(id - 1) * 3 + 1] = out[(id - 1) * 3 + 2] = 0;
    if (!t->in_use[id] || t->lvl[id] < 1) continue;
    int nc = t->ncell[id], kmax = t->dim == 3 ? nc : 1;
    const double *r = CC(t, id, v->rhs), *E = CC(t, id, v->fld), *n = CC(t, id, v->elec);
    double mr = -1.7976931348623157e308, me = mr, mn = mr;
    for (int k = 1; k <= kmax; k++)
      for (int j = 1; j <= nc; j++)
        for (int i = 1; i <= nc; i++) {
          size_t c = t->dim == 2 ? I2(nc, i, j) : I3(nc, i, j, k);
          if (fabs(r[c]) > mr) mr = fabs(r[c]);
          if (E[c] > me) me = E[c];
          if (n[c] > mn) mn = n[c];
        }
    out[(id - 1) * 3 + 0] = t->dr[id] * t->dr[id] * mr;
    out[(id - 1) * 3 + 1] = me;
    out[(id - 1) * 3 + 2] = mn;
  }
}

/* ------------------------------------------------------------------ */
/* Reverse mode of one transport substep of the drivers on a frozen     */
/* mesh: fluxes_koren, a$D_consistent_fluxes, update_solution           */
/* (src/streamer_2d.f90:384-542, src/streamer_3d.f90:396-615,           */
/* src/streamer_cyl.f90:416-570, afivo/src/m_aX_core.f90:818-930). No   */
/* counterpart in the reference, which has no derivatives. Each routine */
/* takes the adjoints of its outputs and adds those of its inputs; every */
/* branch (upwind side, limiter case, flux clamp, table row, dielectric  */
/* rule) is the one the forward pass took on the same data. Pinned by    */
/* central finite differences of the forward routines above             */
/* (tests/test_oracle_adjoint.py).                                      */
/* ------------------------------------------------------------------ */

static void lt_dcol(const otable *t, int col, double x, int li, double lf, double *dval) {
  /* d/dx of lt_col_at at x: the slope of the row interval, 0 where get_loc
   * clamps to the table ends */
  const double *c = t->rc + (size_t)(col - 1) * t->n_rows;
  double frac = (x - t->x_min) * t->inv_dx;
  int raw = (int)ceil(frac);
  (void)lf;
  *dval = (raw < 1 || raw >= t->n_rows) ? 0.0 : (c[li] - c[li - 1]) * t->inv_dx;
}

static void lt_col_adj(const otable *t, int col, int li, double lf, double lam, double *tg) {
  /* adjoint of lf c(li) + (1 - lf) c(li + 1) with respect to the rows */
  if (!tg) return;
  double *g = tg + (size_t)(col - 1) * t->n_rows;
  g[li - 1] += lf * lam;
  g[li] += (1 - lf) * lam;
}

void o_update_solution_adj(otree *t, const o_vars *v, const otable *tbl, double dt,
                           const o_adj *a, double *tgrad) {
  int D = t->dim;
  for (int lvl = t->min_lvl; lvl <= t->max_lvl; lvl++) {
    ivec *lv = &t->leaves[LV(t, lvl)];
    for (int q = 0; q < lv->n; q++) {
      int id = lv->v[q], nc = t->ncell[id];
      double inv_dr = 1 / t->dr[id];
      const double *E = CC(t, id, v->fld), *eps = CC(t, id, v->eps);
      double *le = CC(t, id, a->le), *lp = CC(t, id, a->li), *lE = CC(t, id, a->lE);
      double *lph = v->photoi ? CC(t, id, a->lpho) : NULL;
      const double *F[3];
      double *lF[3];
      for (int d = 0; d < D; d++) {
        F[d] = FC(t, id, v->f_elec, d);
        lF[d] = FC(t, id, a->lF, d);
      }
      int ioff = (IXP(t, id)[0] - 1) * nc;
      int kmax = D == 3 ? nc : 1;
      for (int k = 1; k <= kmax; k++)
        for (int j = 1; j <= nc; j++)
          for (int i = 1; i <= nc; i++) {
            size_t c = D == 2 ? I2(nc, i, j) : I3(nc, i, j, k);
            /* the dielectric rule: n_i' = n_i^* - n_e^*, n_e' = 0 */
            double le_pre = le[c], lp_pre = lp[c];
            if (v->dielectric_rule && eps[c] > 1.0) {
              le_pre = -lp[c];
              lp_pre = lp[c];
            }
            le[c] = le_pre;
            lp[c] = lp_pre;
            double lsrc = le_pre * dt + lp_pre * dt;
            double lsflux = le_pre * dt;
            int li;
            double lf;
            lt_loc(tbl, E[c], &li, &lf);
            double alpha = lt_col_at(tbl, v->t_alpha, li, lf);
            double eta = lt_col_at(tbl, v->t_eta, li, lf);
            size_t fx[2] = {FI(nc, D, 0, i, j, k), FI(nc, D, 0, i + 1, j, k)};
            size_t fy[2] = {FI(nc, D, 1, i, j, k), FI(nc, D, 1, i, j + 1, k)};
            size_t fz[2] = {0, 0};
            double df[3];
            df[0] = F[0][fx[0]] + F[0][fx[1]];
            df[1] = F[1][fy[0]] + F[1][fy[1]];
            if (D == 3) {
              fz[0] = FI(nc, D, 2, i, j, k);
              fz[1] = FI(nc, D, 2, i, j, k + 1);
              df[2] = F[2][fz[0]] + F[2][fz[1]];
            }
            double nrm = o_norm2(D, df);
            /* src = 0.5 |df| (alpha - eta) [+ pho] */
            double lnrm = lsrc * 0.5 * (alpha - eta);
            double lal = lsrc * 0.5 * nrm, let = -lsrc * 0.5 * nrm;
            if (lph) lph[c] += lsrc;
            double da, de;
            lt_dcol(tbl, v->t_alpha, E[c], li, lf, &da);
            lt_dcol(tbl, v->t_eta, E[c], li, lf, &de);
            lE[c] += lal * da + let * de;
            lt_col_adj(tbl, v->t_alpha, li, lf, lal, tgrad);
            lt_col_adj(tbl, v->t_eta, li, lf, let, tgrad);
            double ldf[3] = {0, 0, 0};
            if (nrm > 0)
              for (int d = 0; d < D; d++) ldf[d] = lnrm * df[d] / nrm;
            lF[0][fx[0]] += ldf[0];
            lF[0][fx[1]] += ldf[0];
            lF[1][fy[0]] += ldf[1];
            lF[1][fy[1]] += ldf[1];
            if (D == 3) {
              lF[2][fz[0]] += ldf[2];
              lF[2][fz[1]] += ldf[2];
            }
            double s = lsflux * inv_dr;
            if (D == 2 && t->coord == 1) {
              double den = i + ioff - 0.5;
              double rf1 = (double)(i + ioff - 1) / den, rf2 = (double)(i + ioff) / den;
              lF[0][fx[0]] += rf1 * s;
              lF[0][fx[1]] -= rf2 * s;
            } else {
              lF[0][fx[0]] += s;
              lF[0][fx[1]] -= s;
            }
            lF[1][fy[0]] += s;
            lF[1][fy[1]] -= s;
            if (D == 3) {
              lF[2][fz[0]] += s;
              lF[2][fz[1]] -= s;
            }
          }
    }
  }
}

void o_consistent_fluxes_adj(otree *t, int lF) {
  /* the coarse face values of o_consistent_fluxes are means of fine faces:
   * their adjoints go to those fine faces and are cleared */
  for (int lvl = t->min_lvl; lvl <= t->max_lvl - 1; lvl++) {
    ivec *pa = &t->parents[LV(t, lvl)];
    for (int i = 0; i < pa->n; i++) {
      int id = pa->v[i];
      for (int nb = 1; nb <= t->nnb; nb++) {
        int nb_id = NBP(t, id)[nb - 1];
        if (nb_id <= NO_BOX || has_children(t, nb_id)) continue;
        int nc = t->ncell[id], nch = nc >> 1, d = nb_dim(nb);
        int ii = nb_low(nb) ? 1 : nc + 1, inb = nb_low(nb) ? nc + 1 : 1;
        for (int ic = 0; ic < t->nch / 2; ic++) {
          int ich = ch_adj_nb(t, ic, nb), cid = CHP(t, id)[ich - 1];
          int ioff[3];
          for (int dd = 0; dd < t->dim; dd++) ioff[dd] = nch * ch_dix(ich, dd);
          double *dst = FC(t, nb_id, lF, d), *src = FC(t, cid, lF, d);
          int ta = (d == 0) ? 1 : 0, tb = (d == 2) ? 1 : 2;
          int bmax = t->dim == 3 ? nch : 1;
          for (int b = 1; b <= bmax; b++)
            for (int a = 1; a <= nch; a++) {
              int pd[3], s[3];
              pd[d] = inb; pd[ta] = ioff[ta] + a;
              s[d] = ii; s[ta] = 2 * a - 1;
              if (t->dim == 3) { pd[tb] = ioff[tb] + b; s[tb] = 2 * b - 1; }
              size_t g = FI(nc, t->dim, d, pd[0], pd[1], t->dim == 3 ? pd[2] : 1);
              double lam = dst[g];
              dst[g] = 0;
              if (t->dim == 2) {
                src[FI(nc, 2, d, s[0], s[1], 1)] += 0.5 * lam;
                s[ta] += 1;
                src[FI(nc, 2, d, s[0], s[1], 1)] += 0.5 * lam;
              } else {
                for (int m = 0; m < 4; m++) {
                  int s0[3] = {s[0], s[1], s[2]};
                  s0[ta] += m & 1;
                  s0[tb] += m >> 1;
                  src[FI(nc, 3, d, s0[0], s0[1], s0[2])] += 0.25 * lam;
                }
              }
            }
        }
      }
    }
  }
}

static void gc2_adj(otree *t, int id, int iv, int nb, int x, int b, double lam) {
  /* adjoint of one second-layer ghost value of o_gc2_box (face nb,
   * tangential cell (x, b)): lam goes to the cells it was taken from */
  int nc = t->ncell[id], D = t->dim;
  int nb_id = NBP(t, id)[nb - 1], d = nb_dim(nb);
  int ta = (d == 0) ? 1 : 0, tb = (d == 2) ? 1 : 2;
  if (nb_id != NO_BOX) {
    double *dst = CC(t, nb_id > NO_BOX ? nb_id : id, iv);
    int p[3];
    if (nb_id > NO_BOX) p[d] = nb_low(nb) ? nc - 1 : 2;
    else p[d] = nb_low(nb) ? 2 : nc - 1;
    p[ta] = x;
    if (D == 3) p[tb] = b;
    dst[D == 2 ? I2(nc, p[0], p[1]) : I3(nc, p[0], p[1], p[2])] += lam;
    return;
  }
  int p = t->parent[id], pn = NBP(t, p)[nb - 1], off[3];
  child_offset(t, id, nb, off);
  int pnc = t->ncell[pn];
  double *q = CC(t, pn, iv);
  int ix = nb_hi01(nb) * (nc + 3) - 1;
  int n1 = off[d] + ((ix + 1) >> 1), n2 = n1 + 1 - 2 * (ix & 1);
  int c1 = off[ta] + ((x + 1) >> 1), c2 = c1 + 1 - 2 * (x & 1);
  if (D == 2) {
    int P0[2], P1[2], P2[2];
    P0[d] = n1; P0[ta] = c1;
    P1[d] = n2; P1[ta] = c1;
    P2[d] = n1; P2[ta] = c2;
    q[I2(pnc, P0[0], P0[1])] += 0.5 * lam;
    q[I2(pnc, P1[0], P1[1])] += 0.25 * lam;
    q[I2(pnc, P2[0], P2[1])] += 0.25 * lam;
    return;
  }
  int e1 = off[tb] + ((b + 1) >> 1), e2 = e1 + 1 - 2 * (b & 1);
  int P[4][3];
  for (int m = 0; m < 4; m++) { P[m][d] = n1; P[m][ta] = c1; P[m][tb] = e1; }
  P[1][d] = n2; P[2][ta] = c2; P[3][tb] = e2;
  for (int m = 0; m < 4; m++) q[I3(pnc, P[m][0], P[m][1], P[m][2])] += 0.25 * lam;
}

static void koren_mlim_d(double a, double b, double *da, double *db) {
  /* the derivatives of o_koren_mlim in the case it takes for (a, b) */
  double aa = a * a, ab = a * b;
  if (ab <= 0) { *da = 0; *db = 0; }
  else if (aa <= 0.25 * ab) { *da = 1; *db = 0; }
  else if (aa <= 2.5 * ab) { *da = 1 / 3.0; *db = 1 / 6.0; }
  else { *da = 0; *db = 1; }
}

void o_fluxes_koren_adj(otree *t, const o_vars *v, const otable *tbl, double dt,
                        const o_adj *a, double *tgrad) {
  const double eps0 = 8.8541878176e-12, qe = 1.6022e-19;
  double fac = -0.8 * eps0 / (qe * dt);
  int D = t->dim;
  for (int lvl = t->min_lvl; lvl <= t->max_lvl; lvl++) {
    ivec *lv = &t->leaves[LV(t, lvl)];
    for (int q = 0; q < lv->n; q++) {
      int id = lv->v[q], nc = t->ncell[id];
      double inv_dr = 1 / t->dr[id];
      size_t fl = D == 2 ? (size_t)nc : (size_t)nc * nc;
      double *gc = malloc(sizeof(double) * fl * t->nnb);
      o_gc2_box(t, id, v->elec, gc);
      const double *n = CC(t, id, v->elec), *E = CC(t, id, v->fld);
      double *ln = CC(t, id, a->le), *lE = CC(t, id, a->lE);
      for (int d = 0; d < D; d++) {
        const double *Ff = FC(t, id, v->f_fld, d);
        double *lFe = FC(t, id, a->lF, d), *lFf = FC(t, id, a->lfld, d);
        int n0 = nc + (d == 0), n1 = nc + (d == 1), n2 = D == 3 ? nc + (d == 2) : 1;
        int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
        for (int k = 1; k <= n2; k++)
          for (int j = 1; j <= n1; j++)
            for (int i = 1; i <= n0; i++) {
              int c[3] = {i, j, D == 3 ? k : 0};
#define CI(a0, a1, a2) (D == 2 ? I2(nc, a0, a1) : I3(nc, a0, a1, a2))
              int m1[3] = {c[0], c[1], c[2]}, m2[3] = {c[0], c[1], c[2]}, p1[3] = {c[0], c[1], c[2]};
              m1[d] -= 1;
              m2[d] -= 2;
              p1[d] += 1;
              int pos = c[d];
              size_t ic = CI(c[0], c[1], c[2]), im = CI(m1[0], m1[1], m1[2]);
              size_t fi = FI(nc, D, d, i, j, k);
              double lam = lFe[fi];
              lFe[fi] = 0;
              if (lam == 0) continue;
              double fld_avg = 0.5 * (E[ic] + E[im]);
              int li;
              double lf;
              lt_loc(tbl, fld_avg, &li, &lf);
              double mob = lt_col_at(tbl, v->t_mob, li, lf);
              double dif = lt_col_at(tbl, v->t_dif, li, lf);
              double fld = Ff[fi];
              double vd = -mob * fld;
              double nc0 = n[ic], nm1 = n[im];
              double gradc = nc0 - nm1;
              size_t gpos = (size_t)(c[ta] - 1) + (D == 3 ? (size_t)nc * (c[tb] - 1) : 0);
              int tb_ = D == 3 ? c[tb] : 1;
              /* F = f - dif gradc / dr */
              double ldif = -lam * gradc * inv_dr;
              double lgradc = -lam * dif * inv_dr;
              double lvd = 0, lfld = 0;
              if (vd < 0.0) {
                int ghost = pos == nc + 1;
                double tmp = ghost ? gc[(size_t)(2 * d + 1) * fl + gpos] : n[CI(p1[0], p1[1], p1[2])];
                double gradn = tmp - nc0;
                double lim = o_koren_mlim(gradc, gradn);
                double f = vd * (nc0 - lim);
                if (f < fac * fld) {
                  lfld += lam * fac;
                } else {
                  double da, db;
                  koren_mlim_d(gradc, gradn, &da, &db);
                  lvd = lam * (nc0 - lim);
                  double llim = -lam * vd;
                  ln[ic] += lam * vd;
                  lgradc += llim * da;
                  double lgn = llim * db;
                  ln[ic] -= lgn;
                  if (ghost) gc2_adj(t, id, a->le, 2 * d + 2, c[ta], tb_, lgn);
                  else ln[CI(p1[0], p1[1], p1[2])] += lgn;
                }
              } else {
                int ghost = pos == 1;
                double tmp = ghost ? gc[(size_t)(2 * d) * fl + gpos] : n[CI(m2[0], m2[1], m2[2])];
                double gradp = nm1 - tmp;
                double lim = o_koren_mlim(gradc, gradp);
                double f = vd * (nm1 + lim);
                if (f > fac * fld) {
                  lfld += lam * fac;
                } else {
                  double da, db;
                  koren_mlim_d(gradc, gradp, &da, &db);
                  lvd = lam * (nm1 + lim);
                  double llim = lam * vd;
                  ln[im] += lam * vd;
                  lgradc += llim * da;
                  double lgp = llim * db;
                  ln[im] += lgp;
                  if (ghost) gc2_adj(t, id, a->le, 2 * d + 1, c[ta], tb_, -lgp);
                  else ln[CI(m2[0], m2[1], m2[2])] -= lgp;
                }
              }
              ln[ic] += lgradc;
              ln[im] -= lgradc;
              /* vd = -mob fld */
              double lmob = -lvd * fld;
              lfld += -lvd * mob;
              lFf[fi] += lfld;
              double dm, dd;
              lt_dcol(tbl, v->t_mob, fld_avg, li, lf, &dm);
              lt_dcol(tbl, v->t_dif, fld_avg, li, lf, &dd);
              double lx = lmob * dm + ldif * dd;
              lE[ic] += 0.5 * lx;
              lE[im] += 0.5 * lx;
              lt_col_adj(tbl, v->t_mob, li, lf, lmob, tgrad);
              lt_col_adj(tbl, v->t_dif, li, lf, ldif, tgrad);
#undef CI
            }
      }
      free(gc);
    }
  }
}

void o_lt_set_rows(otable *t, const double *rc) {
  /* the table's rows_cols(n_rows, n_cols) as given (test perturbations) */
  memcpy(t->rc, rc, sizeof(double) * (size_t)t->n_rows * t->n_cols);
}

/* ------------------------------------------------------------------ */
/* reverse mode of the rest of a time step and of the FAS FMG          */
/* (no counterpart in the reference, which computes no derivatives)   */
/* ------------------------------------------------------------------ */
/* Each routine takes the adjoints of the values its forward routine writes
 * (and clears them) and adds those of the values it reads, one loop per
 * forward loop, in plain sequential code. The GPU's reverse kernels scatter
 * the same terms with atomics; these sums come in loop order. */

void o_tree_copy_cc_adj(otree *t, int lfrom, int lto) {
  /* a$D_tree_copy_cc (m_aX_utils.f90:432-453), to = from on whole boxes */
  for (int lvl = t->min_lvl; lvl <= t->max_lvl; lvl++) {
    ivec *ids = &t->ids[LV(t, lvl)];
    for (int q = 0; q < ids->n; q++) {
      int id = ids->v[q], n = blen(t, id);
      double *f = CC(t, id, lfrom), *o = CC(t, id, lto);
      for (int m = 0; m < n; m++) {
        f[m] += o[m];
        o[m] = 0;
      }
    }
  }
}

void o_average_dens_adj(otree *t, int le, int leo, int li, int lio) {
  /* average_dens (src/streamer_2d.f90:492-496), n = 0.5 (n + n_old) */
  for (int lvl = t->min_lvl; lvl <= t->max_lvl; lvl++) {
    ivec *ids = &t->ids[LV(t, lvl)];
    for (int q = 0; q < ids->n; q++) {
      int id = ids->v[q], n = blen(t, id);
      double *e = CC(t, id, le), *eo = CC(t, id, leo), *p = CC(t, id, li), *po = CC(t, id, lio);
      for (int m = 0; m < n; m++) {
        eo[m] += 0.5 * e[m];
        e[m] = 0.5 * e[m];
        po[m] += 0.5 * p[m];
        p[m] = 0.5 * p[m];
      }
    }
  }
}

void o_space_charge_adj(otree *t, double fac, int lr, int le, int li) {
  /* rhs = fac (n_e - n_i) on leaf boxes (src/streamer_2d.f90:318-329) */
  for (int lvl = 1; lvl <= t->max_lvl; lvl++) {
    ivec *lv = &t->leaves[LV(t, lvl)];
    for (int q = 0; q < lv->n; q++) {
      int id = lv->v[q], n = blen(t, id);
      double *r = CC(t, id, lr), *e = CC(t, id, le), *p = CC(t, id, li);
      for (int m = 0; m < n; m++) {
        e[m] += fac * r[m];
        p[m] -= fac * r[m];
        r[m] = 0;
      }
    }
  }
}

static void restrict_box_adj(otree *t, int c_id, int p_id, int il) {
  /* a$D_restrict_box (m_aX_restrict.f90:67-142) transposed */
  int nc = t->ncell[c_id], hnc = nc >> 1, pnc = t->ncell[p_id], off[3];
  child_offset(t, c_id, 0, off);
  double *c = CC(t, c_id, il), *p = CC(t, p_id, il);
  if (t->dim == 2) {
    double dr16 = 0.0625 * t->dr[p_id];
    for (int j = 1; j <= hnc; j++) {
      int jc = off[1] + j, jf = 2 * j - 1;
      for (int i = 1; i <= hnc; i++) {
        int ic = off[0] + i, fi = 2 * i - 1;
        double lam = p[I2(pnc, ic, jc)];
        p[I2(pnc, ic, jc)] = 0;
        if (t->coord == 1) {
          double r = RMP(t, p_id)[0] + (i - 0.5) * t->dr[p_id];
          double rfac = dr16 / r;
          c[I2(nc, fi, jf)] += (0.25 - rfac) * lam;
          c[I2(nc, fi, jf + 1)] += (0.25 - rfac) * lam;
          c[I2(nc, fi + 1, jf)] += (0.25 + rfac) * lam;
          c[I2(nc, fi + 1, jf + 1)] += (0.25 + rfac) * lam;
        } else {
          c[I2(nc, fi, jf)] += 0.25 * lam;
          c[I2(nc, fi + 1, jf)] += 0.25 * lam;
          c[I2(nc, fi, jf + 1)] += 0.25 * lam;
          c[I2(nc, fi + 1, jf + 1)] += 0.25 * lam;
        }
      }
    }
  } else {
    for (int k = 1; k <= hnc; k++)
      for (int j = 1; j <= hnc; j++)
        for (int i = 1; i <= hnc; i++) {
          size_t pc = I3(pnc, off[0] + i, off[1] + j, off[2] + k);
          double lam = p[pc];
          p[pc] = 0;
          for (int dk = 0; dk < 2; dk++)
            for (int dj = 0; dj < 2; dj++)
              for (int di = 0; di < 2; di++)
                c[I3(nc, 2 * i - 1 + di, 2 * j - 1 + dj, 2 * k - 1 + dk)] += 0.125 * lam;
        }
  }
}

void o_restrict_tree_adj(otree *t, int il) {
  /* a$D_restrict_tree (m_aX_restrict.f90:54-64) transposed: coarsest first */
  for (int lvl = t->min_lvl; lvl <= t->max_lvl - 1; lvl++) {
    ivec *pa = &t->parents[LV(t, lvl)];
    for (int q = 0; q < pa->n; q++) {
      int id = pa->v[q];
      for (int c = 0; c < t->nch; c++) {
        int cid = CHP(t, id)[c];
        if (cid != NO_BOX) restrict_box_adj(t, cid, id, il);
      }
    }
  }
}

/* a$D_gc_interp[_lim] (m_aX_gc.f90:90-295) on face nb, transposed; the
 * limiter's case is recomputed from the primal iv */
static void gc_interp_adj(otree *t, int id, int nb, int iv, int il, int lim) {
  const double sixth = 1 / 6.0, third = 1 / 3.0;
  int nc = t->ncell[id], p = t->parent[id], pn = NBP(t, p)[nb - 1], off[3];
  child_offset(t, id, nb, off);
  int ix, ixf, ixc;
  if (nb_low(nb)) { ix = 0; ixf = 1; ixc = nc; }
  else { ix = nc + 1; ixf = nc; ixc = 1; }
  double *a = CC(t, id, iv), *q = CC(t, pn, iv);
  double *la = CC(t, id, il), *lq = CC(t, pn, il);
  int pnc = t->ncell[pn], d = nb_dim(nb);
  if (t->dim == 2) {
    for (int x = 1; x <= nc; x++) {
      int o = (d == 0) ? off[1] : off[0];
      int c1 = o + ((x + 1) >> 1), c2 = c1 + 1 - 2 * (x & 1);
      size_t q1, q2, g, f;
      if (d == 0) {
        q1 = I2(pnc, ixc, c1); q2 = I2(pnc, ixc, c2); g = I2(nc, ix, x); f = I2(nc, ixf, x);
      } else {
        q1 = I2(pnc, c1, ixc); q2 = I2(pnc, c2, ixc); g = I2(nc, x, ix); f = I2(nc, x, ixf);
      }
      double lam = la[g];
      la[g] = 0;
      double v = 0.5 * q[q1] + sixth * q[q2] + third * a[f];
      if (lim && v > 2 * q[q1]) {
        lq[q1] += 2 * lam;
      } else {
        lq[q1] += 0.5 * lam;
        lq[q2] += sixth * lam;
        la[f] += third * lam;
      }
    }
  } else {
    for (int y = 1; y <= nc; y++)
      for (int x = 1; x <= nc; x++) {
        size_t q1, q2, q3, g, f;
        if (d == 0) {
          int k = y, j = x;
          int kc1 = off[2] + ((k + 1) >> 1), kc2 = kc1 + 1 - 2 * (k & 1);
          int jc1 = off[1] + ((j + 1) >> 1), jc2 = jc1 + 1 - 2 * (j & 1);
          q1 = I3(pnc, ixc, jc1, kc1); q2 = I3(pnc, ixc, jc2, kc1); q3 = I3(pnc, ixc, jc1, kc2);
          g = I3(nc, ix, j, k); f = I3(nc, ixf, j, k);
        } else if (d == 1) {
          int k = y, i = x;
          int kc1 = off[2] + ((k + 1) >> 1), kc2 = kc1 + 1 - 2 * (k & 1);
          int ic1 = off[0] + ((i + 1) >> 1), ic2 = ic1 + 1 - 2 * (i & 1);
          q1 = I3(pnc, ic1, ixc, kc1); q2 = I3(pnc, ic2, ixc, kc1); q3 = I3(pnc, ic1, ixc, kc2);
          g = I3(nc, i, ix, k); f = I3(nc, i, ixf, k);
        } else {
          int j = y, i = x;
          int jc1 = off[1] + ((j + 1) >> 1), jc2 = jc1 + 1 - 2 * (j & 1);
          int ic1 = off[0] + ((i + 1) >> 1), ic2 = ic1 + 1 - 2 * (i & 1);
          q1 = I3(pnc, ic1, jc1, ixc); q2 = I3(pnc, ic1, jc2, ixc); q3 = I3(pnc, ic2, jc1, ixc);
          g = I3(nc, i, j, ix); f = I3(nc, i, j, ixf);
        }
        double lam = la[g];
        la[g] = 0;
        double v = third * q[q1] + sixth * q[q2] + sixth * q[q3] + third * a[f];
        if (lim && v > 2 * q[q1]) {
          lq[q1] += 2 * lam;
        } else {
          lq[q1] += third * lam;
          lq[q2] += sixth * lam;
          lq[q3] += sixth * lam;
          la[f] += third * lam;
        }
      }
  }
}

/* sides_from_nb (m_aX_gc.f90:368-402) transposed */
static void sides_from_nb_adj(otree *t, int id, int nb_id, int nb, int il) {
  int nc = t->ncell[id];
  double *a = CC(t, id, il), *o = CC(t, nb_id, il);
  int d = nb_dim(nb);
  int src = nb_low(nb) ? nc : 1, dst = nb_low(nb) ? 0 : nc + 1;
  int bmax = t->dim == 3 ? nc : 1;
  for (int b = 1; b <= bmax; b++)
    for (int x = 1; x <= nc; x++) {
      int p[3], q[3];
      int ta = (d == 0) ? 1 : 0, tb = (d == 2) ? 1 : 2;
      p[d] = dst; q[d] = src;
      p[ta] = q[ta] = x;
      if (t->dim == 3) p[tb] = q[tb] = b;
      size_t pi = t->dim == 2 ? I2(nc, p[0], p[1]) : I3(nc, p[0], p[1], p[2]);
      size_t qi = t->dim == 2 ? I2(nc, q[0], q[1]) : I3(nc, q[0], q[1], q[2]);
      o[qi] += a[pi];
      a[pi] = 0;
    }
}

void o_gc_tree_adj(otree *t, int iv, int il, int rb, int bc) {
  /* a$D_gc_tree (m_aX_gc.f90:28-46) with a$D_gc_interp (rb 1) or
   * a$D_gc_interp_lim (rb 2) and a$D_gc_neumann (bc 1) or a$D_gc_dirichlet
   * (bc 2), transposed: levels and faces in the reverse order */
  for (int lvl = t->max_lvl; lvl >= t->min_lvl; lvl--) {
    ivec *ids = &t->ids[LV(t, lvl)];
    for (int q = ids->n - 1; q >= 0; q--) {
      int id = ids->v[q];
      for (int nb = t->nnb; nb >= 1; nb--) {
        int nb_id = NBP(t, id)[nb - 1];
        if (nb_id > NO_BOX) {
          sides_from_nb_adj(t, id, nb_id, nb, il);
        } else if (nb_id == NO_BOX) {
          gc_interp_adj(t, id, nb, iv, il, rb == 2);
        } else {
          double *a = CC(t, id, il);
          FACE_LOOP_BEGIN(t, id, nb)
          a[c] += bc == 1 ? a[g] : -a[g];
          a[g] = 0;
          FACE_LOOP_END
        }
      }
    }
  }
}

void o_fld_from_pot_adj(otree *t, const o_vars *v, int lE, int lfld, int lphi) {
  /* fld_from_pot (src/streamer_2d.f90:346-381, src/streamer_3d.f90:348-395)
   * transposed on every box: the interior |E| adjoints to the faces
   * (d|E|/dF = (F_lo + F_hi) / (4 |E|), zero where |E| = 0), then every
   * face adjoint to the potential of its two cells */
  for (int lvl = t->min_lvl; lvl <= t->max_lvl; lvl++) {
    ivec *ids = &t->ids[LV(t, lvl)];
    for (int q = 0; q < ids->n; q++) {
      int id = ids->v[q], nc = t->ncell[id], D = t->dim;
      double inv_dr = 1 / t->dr[id];
      const double *e = CC(t, id, v->eps), *E = CC(t, id, v->fld);
      double *LE = CC(t, id, lE), *LP = CC(t, id, lphi);
      const double *F[3];
      double *LF[3];
      for (int d = 0; d < D; d++) {
        F[d] = FC(t, id, v->f_fld, d);
        LF[d] = FC(t, id, lfld, d);
      }
      int kmax = D == 3 ? nc : 1;
      for (int k = 1; k <= kmax; k++)
        for (int j = 1; j <= nc; j++)
          for (int i = 1; i <= nc; i++) {
            size_t c = D == 2 ? I2(nc, i, j) : I3(nc, i, j, k);
            double le = LE[c], ev = E[c];
            LE[c] = 0;
            if (le == 0 || !(ev > 0)) continue;
            for (int d = 0; d < D; d++) {
              int h[3] = {i, j, k};
              h[d] += 1;
              size_t f0 = FI(nc, D, d, i, j, k), f1 = FI(nc, D, d, h[0], h[1], h[2]);
              double s = F[d][f0] + F[d][f1];
              LF[d][f0] += le * (0.25 * s / ev);
              LF[d][f1] += le * (0.25 * s / ev);
            }
          }
      for (int d = 0; d < D; d++) {
        int n0 = nc + (d == 0), n1 = nc + (d == 1), n2 = D == 3 ? nc + (d == 2) : 1;
        for (int k = 1; k <= n2; k++)
          for (int j = 1; j <= n1; j++)
            for (int i = 1; i <= n0; i++) {
              int a[3] = {i, j, D == 3 ? k : 0}, b[3] = {i, j, D == 3 ? k : 0};
              a[d] -= 1;
              size_t ia = D == 2 ? I2(nc, a[0], a[1]) : I3(nc, a[0], a[1], a[2]);
              size_t ib = D == 2 ? I2(nc, b[0], b[1]) : I3(nc, b[0], b[1], b[2]);
              int fpos = (d == 0) ? i : (d == 1) ? j : k;
              double kf;
              if (fpos == 1) kf = 2 * inv_dr * e[ia] / (e[ib] + e[ia]);
              else if (fpos == nc + 1) kf = 2 * inv_dr * e[ib] / (e[ib] + e[ia]);
              else kf = inv_dr;
              size_t fi = FI(nc, D, d, i, j, k);
              double lam = LF[d][fi];
              LF[d][fi] = 0;
              LP[ia] += kf * lam;
              LP[ib] -= kf * lam;
            }
      }
    }
  }
}

/* --- the FAS FMG (m_aX_mg.f90:164-255) transposed: Cartesian or        */
/* cylindrical Laplacian, no permittivity, no Helmholtz term              */

typedef struct {
  int lp, lt, lr, lw;
  double *vbar;
} o_adjmg;

/* a$D_gc_box with mg$D_sides_rb and the multigrid face rules, transposed:
 * faces in the reverse order; on a y face with the whole-row rule its
 * corner ghosts (written last, from the x ghosts) go first */
static void gc_box_mg_adj(otree *t, int id, const o_adjmg *m) {
  int nc = t->ncell[id], D = t->dim;
  double *a = CC(t, id, m->lp);
  for (int nb = t->nnb; nb >= 1; nb--) {
    int nb_id = NBP(t, id)[nb - 1];
    if (nb_id > NO_BOX) {
      sides_from_nb_adj(t, id, nb_id, nb, m->lp);
    } else if (nb_id == NO_BOX) {
      /* 0.5 times the parent's cell (a$D_gc_prolong0) plus an
       * extrapolation from the box's own cells (mg$D_sides_rb) */
      int d = nb_dim(nb), dix = nb_low(nb) ? 1 : -1;
      int p = t->parent[id], pnc = t->ncell[p], off[3];
      child_offset(t, id, 0, off);
      double *lpp = CC(t, p, m->lp);
      FACE_LOOP_BEGIN(t, id, nb)
      double lam = a[g];
      a[g] = 0;
      if (lam != 0) {
        int gi[3] = {q_[0], q_[1], D_ == 3 ? q_[2] : 0};
        int pc[3] = {0, 0, 0};
        for (int e2 = 0; e2 < D; e2++) pc[e2] = off[e2] + ((gi[e2] + 1) >> 1);
        lpp[D == 2 ? I2(pnc, pc[0], pc[1]) : I3(pnc, pc[0], pc[1], pc[2])] += 0.5 * lam;
        a[c] += (D == 2 ? 1.0 : 1.25) * lam;
        int e1[3] = {r_[0], r_[1], D_ == 3 ? r_[2] : 0};
        e1[d] += dix;
        a[D == 2 ? I2(nc, e1[0], e1[1]) : I3(nc, e1[0], e1[1], e1[2])] -= 0.25 * lam;
        int ta = (d == 0) ? 1 : 0, tb = (d == 2) ? 1 : 2;
        int e2[3] = {r_[0], r_[1], D_ == 3 ? r_[2] : 0};
        e2[ta] += -1 + 2 * (e2[ta] & 1);
        a[D == 2 ? I2(nc, e2[0], e2[1]) : I3(nc, e2[0], e2[1], e2[2])] -= 0.25 * lam;
        if (D == 3) {
          int e3[3] = {r_[0], r_[1], r_[2]};
          e3[tb] += -1 + 2 * (e3[tb] & 1);
          a[I3(nc, e3[0], e3[1], e3[2])] -= 0.25 * lam;
        }
      }
      FACE_LOOP_END
    } else {
      int kind = t->bc_kind[nb - 1];
      if (kind == 3 && D == 2 && nb_dim(nb) == 1) {
        int gy = nb_low(nb) ? 0 : nc + 1, cy = nb_low(nb) ? 1 : nc;
        for (int s = 0; s < 2; s++) {
          int gx = s ? nc + 1 : 0;
          double lam = a[I2(nc, gx, gy)];
          a[I2(nc, gx, gy)] = 0;
          a[I2(nc, gx, cy)] -= lam;
          m->vbar[nb - 1] += 2 * lam;
        }
      }
      FACE_LOOP_BEGIN(t, id, nb)
      double lam = a[g];
      a[g] = 0;
      if (kind == 0) a[c] += lam;
      else if (kind == 1) a[c] -= lam;
      else {
        a[c] -= lam;
        m->vbar[nb - 1] += 2 * lam;
      }
      FACE_LOOP_END
    }
  }
}

static void adj_gc_level(otree *t, int lvl, const o_adjmg *m) {
  ivec *ids = &t->ids[LV(t, lvl)];
  for (int q = ids->n - 1; q >= 0; q--) gc_box_mg_adj(t, ids->v[q], m);
}

/* the neighbours of interior cell (i, j, k), in the order i-1, i+1, j-1,
 * j+1, k-1, k+1 of the reference's u(:) and a(:) */
static int nb_index(int D, int nc, int i, int j, int k, size_t *nbi) {
  if (D == 2) {
    nbi[0] = I2(nc, i - 1, j); nbi[1] = I2(nc, i + 1, j);
    nbi[2] = I2(nc, i, j - 1); nbi[3] = I2(nc, i, j + 1);
  } else {
    nbi[0] = I3(nc, i - 1, j, k); nbi[1] = I3(nc, i + 1, j, k);
    nbi[2] = I3(nc, i, j - 1, k); nbi[3] = I3(nc, i, j + 1, k);
    nbi[4] = I3(nc, i, j, k - 1); nbi[5] = I3(nc, i, j, k + 1);
  }
  return 2 * D;
}

/* one half-sweep (mg$D_box_gsrb_lpl, mg2_box_gsrb_clpl; for a box tagged
 * with a permittivity gsrb_lpld above) transposed */
static void gsrb_box_adj(otree *t, int id, int cntr, const o_adjmg *m) {
  int nc = t->ncell[id], D = t->dim;
  int cyl = D == 2 && t->coord == 1 && t->auto_ops;
  double dx2 = t->dr[id] * t->dr[id];
  double *L = CC(t, id, m->lp), *R = CC(t, id, m->lr);
  const double *e = box_tag(t, id) == TAG_NORMAL ? NULL : CC(t, id, t->i_eps);
  /* 1 / 2D, or with a Helmholtz term 1 / (2D + helm dx2) */
  double w = t->helm != 0 ? 1 / (2 * D + t->helm * dx2) : (D == 2 ? 0.25 : 1 / 6.0);
  int kmax = D == 3 ? nc : 1;
  int ioff = cyl ? (IXP(t, id)[0] - 1) * nc : 0;
  for (int k = 1; k <= kmax; k++)
    for (int j = 1; j <= nc; j++) {
      int i0 = 2 - ((cntr ^ (D == 3 ? k + j : j)) & 1);
      for (int i = i0; i <= nc; i += 2) {
        size_t c = D == 2 ? I2(nc, i, j) : I3(nc, i, j, k);
        double y = L[c];
        L[c] = 0;
        if (y == 0) continue;
        if (e) {
          /* p = (sum c_m rf_m u_m - dx2 rhs) / sum c_m rf_m */
          size_t nbi[6];
          int nn = nb_index(D, nc, i, j, k, nbi);
          double a0 = e[c], cm[6], s2 = 0;
          for (int q = 0; q < nn; q++) {
            cm[q] = 2 * a0 * e[nbi[q]] / (a0 + e[nbi[q]]);
            if (D == 2 && t->coord == 1 && q < 2) {
              double den = i + (IXP(t, id)[0] - 1) * nc - 0.5;
              cm[q] = cm[q] * ((double)(i + (IXP(t, id)[0] - 1) * nc - 1 + q) / den);
            }
            s2 = s2 + cm[q];
          }
          for (int q = 0; q < nn; q++) L[nbi[q]] += cm[q] / s2 * y;
          R[c] += -dx2 / s2 * y;
          continue;
        }
        double wxm = w, wxp = w;
        if (cyl) {
          double den = i + ioff - 0.5;
          wxm = w * ((double)(i + ioff - 1) / den);
          wxp = w * ((double)(i + ioff) / den);
        }
        if (D == 2) {
          L[I2(nc, i - 1, j)] += wxm * y;
          L[I2(nc, i + 1, j)] += wxp * y;
          L[I2(nc, i, j - 1)] += w * y;
          L[I2(nc, i, j + 1)] += w * y;
        } else {
          L[I3(nc, i - 1, j, k)] += w * y;
          L[I3(nc, i + 1, j, k)] += w * y;
          L[I3(nc, i, j - 1, k)] += w * y;
          L[I3(nc, i, j + 1, k)] += w * y;
          L[I3(nc, i, j, k - 1)] += w * y;
          L[I3(nc, i, j, k + 1)] += w * y;
        }
        R[c] += -w * dx2 * y;
      }
    }
}

static void adj_gsrb_boxes(otree *t, int lvl, int n_cycle, const o_adjmg *m) {
  ivec *ids = &t->ids[LV(t, lvl)];
  for (int it = 2 * n_cycle; it >= 1; it--) {
    adj_gc_level(t, lvl, m);
    for (int q = ids->n - 1; q >= 0; q--) gsrb_box_adj(t, ids->v[q], it, m);
  }
}

/* dst += s A^T src on box id (A the box Laplacian, or box_lpld for a box
 * tagged with a permittivity); src cleared if asked */
static void op_box_adj(otree *t, int id, int isrc, double s, int idst, int clear) {
  int nc = t->ncell[id], D = t->dim;
  int cyl = D == 2 && t->coord == 1 && t->auto_ops;
  double inv = 1 / (t->dr[id] * t->dr[id]);
  double *S = CC(t, id, isrc), *Dp = CC(t, id, idst);
  const double *e = box_tag(t, id) == TAG_NORMAL ? NULL : CC(t, id, t->i_eps);
  int kmax = D == 3 ? nc : 1;
  int ioff = cyl ? (IXP(t, id)[0] - 1) * nc : 0;
  for (int k = 1; k <= kmax; k++)
    for (int j = 1; j <= nc; j++)
      for (int i = 1; i <= nc; i++) {
        size_t c = D == 2 ? I2(nc, i, j) : I3(nc, i, j, k);
        double sb = s * S[c];
        if (clear) S[c] = 0;
        if (sb == 0) continue;
        if (e) {
          /* o = 2 inv sum_m rf_m a0 a_m / (a0 + a_m) (u_m - u0) */
          size_t nbi[6];
          int nn = nb_index(D, nc, i, j, k, nbi);
          double a0 = e[c], sum = 0;
          for (int q = 0; q < nn; q++) {
            double cm = a0 * e[nbi[q]] / (a0 + e[nbi[q]]);
            if (D == 2 && t->coord == 1 && q < 2) {
              double den = i + (IXP(t, id)[0] - 1) * nc - 0.5;
              cm = ((double)(i + (IXP(t, id)[0] - 1) * nc - 1 + q) / den) * cm;
            }
            Dp[nbi[q]] += 2 * inv * cm * sb;
            sum = sum + cm;
          }
          Dp[c] += -2 * inv * sum * sb;
          continue;
        }
        double wxm = inv, wxp = inv;
        if (cyl) {
          double den = i + ioff - 0.5;
          wxm = ((double)(i + ioff - 1) / den) * inv;
          wxp = ((double)(i + ioff) / den) * inv;
        }
        if (D == 2) {
          Dp[I2(nc, i - 1, j)] += wxm * sb;
          Dp[I2(nc, i + 1, j)] += wxp * sb;
          Dp[I2(nc, i, j - 1)] += inv * sb;
          Dp[I2(nc, i, j + 1)] += inv * sb;
          Dp[c] += -4.0 * inv * sb;
        } else {
          Dp[I3(nc, i - 1, j, k)] += inv * sb;
          Dp[I3(nc, i + 1, j, k)] += inv * sb;
          Dp[I3(nc, i, j - 1, k)] += inv * sb;
          Dp[I3(nc, i, j + 1, k)] += inv * sb;
          Dp[I3(nc, i, j, k - 1)] += inv * sb;
          Dp[I3(nc, i, j, k + 1)] += inv * sb;
          Dp[c] += -6.0 * inv * sb;
        }
        if (t->helm != 0) Dp[c] += -t->helm * sb;  /* lpl(u) - helm u */
      }
}

/* mg$D_box_rstr_lpl of child id transposed: the parent's adjoint in ip to
 * the child's cells in idst (added or assigned), the parent's cleared */
static void rstr_box_adj(otree *t, int id, int ip, int idst, int assign) {
  int p_id = t->parent[id], nc = t->ncell[id], hnc = nc >> 1, pnc = t->ncell[p_id], off[3];
  child_offset(t, id, 0, off);
  double *P = CC(t, p_id, ip), *C = CC(t, id, idst);
  int D = t->dim;
  double w = D == 2 ? 0.25 : 0.125;
  int kmax = D == 3 ? hnc : 1;
  for (int k = 1; k <= kmax; k++)
    for (int j = 1; j <= hnc; j++)
      for (int i = 1; i <= hnc; i++) {
        size_t pc = D == 2 ? I2(pnc, off[0] + i, off[1] + j)
                           : I3(pnc, off[0] + i, off[1] + j, off[2] + k);
        double v = w * P[pc];
        P[pc] = 0;
        for (int f = 0; f < (D == 2 ? 4 : 8); f++) {
          int fi = 2 * i - 1 + (f & 1), fj = 2 * j - 1 + ((f >> 1) & 1);
          int fk = D == 3 ? 2 * k - 1 + (f >> 2) : 0;
          size_t c = D == 2 ? I2(nc, fi, fj) : I3(nc, fi, fj, fk);
          C[c] = assign ? v : C[c] + v;
        }
      }
}

/* mg$D_box_corr_lpl of child id transposed: the child's potential adjoint
 * (unchanged) into the parent's tmp adjoint with the prolongation weights */
static void corr_box_adj(otree *t, int id, const o_adjmg *m) {
  int p_id = t->parent[id], nc = t->ncell[id], pnc = t->ncell[p_id], off[3], D = t->dim;
  child_offset(t, id, 0, off);
  const double *F = CC(t, id, m->lp);
  double *Q = CC(t, p_id, m->lt);
  /* mg$D_box_corr_lpld for a child tagged with a permittivity: the parent's
   * eps weights the prolongation (o_box_corr above) */
  const double *e = box_tag(t, id) == TAG_NORMAL ? NULL : CC(t, p_id, t->i_eps);
  int kmax = D == 3 ? nc : 1;
  for (int k = 1; k <= kmax; k++)
    for (int j = 1; j <= nc; j++)
      for (int i = 1; i <= nc; i++) {
        double fb = F[D == 2 ? I2(nc, i, j) : I3(nc, i, j, k)];
        if (fb == 0) continue;
        int ic1 = off[0] + ((i + 1) >> 1), ic2 = ic1 + 1 - 2 * (i & 1);
        int jc1 = off[1] + ((j + 1) >> 1), jc2 = jc1 + 1 - 2 * (j & 1);
        if (e) {
          int kc1 = D == 3 ? off[2] + ((k + 1) >> 1) : 0, kc2 = D == 3 ? kc1 + 1 - 2 * (k & 1) : 0;
          size_t p0 = D == 2 ? I2(pnc, ic1, jc1) : I3(pnc, ic1, jc1, kc1);
          size_t pm[3];
          if (D == 2) {
            pm[0] = I2(pnc, ic2, jc1); pm[1] = I2(pnc, ic1, jc2);
          } else {
            pm[0] = I3(pnc, ic2, jc1, kc1); pm[1] = I3(pnc, ic1, jc2, kc1);
            pm[2] = I3(pnc, ic1, jc1, kc2);
          }
          double a0 = e[p0], w0 = 0;
          for (int q = 0; q < D; q++) {
            double am = e[pm[q]];
            /* 2D: 1/2 (a0 u0 + a_m u_m) / (a0 + a_m); 3D: 1/3 (a0 u0 + a_m
             * (3/2 u_m - 1/2 u0)) / (a0 + a_m), summed over m */
            Q[pm[q]] += 0.5 * am / (a0 + am) * fb;
            w0 += (D == 2 ? 0.5 * a0 : (a0 - 0.5 * am) / 3.0) / (a0 + am);
          }
          Q[p0] += w0 * fb;
          continue;
        }
        if (D == 2) {
          Q[I2(pnc, ic1, jc1)] += 0.5 * fb;
          Q[I2(pnc, ic2, jc1)] += 0.25 * fb;
          Q[I2(pnc, ic1, jc2)] += 0.25 * fb;
        } else {
          int kc1 = off[2] + ((k + 1) >> 1), kc2 = kc1 + 1 - 2 * (k & 1);
          Q[I3(pnc, ic1, jc1, kc1)] += 0.25 * fb;
          Q[I3(pnc, ic2, jc1, kc1)] += 0.25 * fb;
          Q[I3(pnc, ic1, jc2, kc1)] += 0.25 * fb;
          Q[I3(pnc, ic1, jc1, kc2)] += 0.25 * fb;
        }
      }
}

/* whole-box transposes: mode 0 tmp = phi; mode 1 tmp = phi - tmp; mode 2
 * rhs += tmp */
static void whole_adj(otree *t, int id, int mode, const o_adjmg *m) {
  int n = blen(t, id);
  double *Lp = CC(t, id, m->lp), *Lt = CC(t, id, m->lt), *Lr = CC(t, id, m->lr);
  for (int q = 0; q < n; q++) {
    if (mode == 2) {
      Lt[q] += Lr[q];
    } else {
      double tb = Lt[q];
      Lp[q] += tb;
      Lt[q] = mode == 0 ? 0.0 : -tb;
    }
  }
}

/* dst += src on the interior of box id; src cleared there if asked */
static void int_add_adj(otree *t, int id, int isrc, int idst, int clear) {
  int nc = t->ncell[id], D = t->dim, kmax = D == 3 ? nc : 1;
  double *S = CC(t, id, isrc), *Dp = CC(t, id, idst);
  for (int k = 1; k <= kmax; k++)
    for (int j = 1; j <= nc; j++)
      for (int i = 1; i <= nc; i++) {
        size_t c = D == 2 ? I2(nc, i, j) : I3(nc, i, j, k);
        Dp[c] += S[c];
        if (clear) S[c] = 0;
      }
}

static void adj_correct_children(otree *t, int lvl, const o_adjmg *m) {
  /* correct_children over the parents of lvl (m_aX_mg.f90:374-401) */
  ivec *pa = &t->parents[LV(t, lvl)];
  for (int q = pa->n - 1; q >= 0; q--) {
    int id = pa->v[q];
    for (int c = t->nch - 1; c >= 0; c--) {
      int cid = CHP(t, id)[c];
      if (cid != NO_BOX) corr_box_adj(t, cid, m);
    }
    whole_adj(t, id, 1, m);
  }
}

static void adj_update_coarse(otree *t, int lvl, int keep_res, const o_adjmg *m) {
  /* update_coarse (keep_res 0, the fine tmp restored; m_aX_mg.f90:430-486)
   * and set_coarse_phi_rhs (keep_res 1; :490-523) transposed */
  ivec *pa = &t->parents[LV(t, lvl - 1)];
  for (int q = pa->n - 1; q >= 0; q--) {
    int id = pa->v[q];
    if (!keep_res) whole_adj(t, id, 0, m);           /* tmp = phi */
    whole_adj(t, id, 2, m);                          /* rhs += tmp */
    op_box_adj(t, id, m->lr, 1.0, m->lp, 1);         /* rhs = A phi */
  }
  adj_gc_level(t, lvl - 1, m);
  ivec *ids = &t->ids[LV(t, lvl)];
  for (int q = ids->n - 1; q >= 0; q--) {
    int id = ids->v[q];
    rstr_box_adj(t, id, m->lp, m->lp, 0);            /* phi_p = R phi */
    if (keep_res) {
      rstr_box_adj(t, id, m->lt, m->lt, 0);          /* tmp_p = R tmp */
      op_box_adj(t, id, m->lt, -1.0, m->lp, 0);      /* tmp = rhs - A phi */
      int_add_adj(t, id, m->lt, m->lr, 1);
    } else {
      rstr_box_adj(t, id, m->lt, m->lw, 1);          /* tmp_p = R (rhs - A phi) */
      op_box_adj(t, id, m->lw, -1.0, m->lp, 0);
      int_add_adj(t, id, m->lw, m->lr, 0);
    }
  }
  if (keep_res && lvl == t->max_lvl) adj_gc_level(t, lvl, m);
}

static void adj_vcycle(otree *t, int max_lvl, const o_adjmg *m) {
  int min_lvl = t->min_lvl;
  for (int lvl = max_lvl; lvl >= min_lvl + 1; lvl--) {
    adj_gsrb_boxes(t, lvl, t->n_up, m);
    adj_gc_level(t, lvl, m);
    adj_correct_children(t, lvl - 1, m);
  }
  adj_gsrb_boxes(t, min_lvl, t->n_base, m);
  for (int lvl = min_lvl + 1; lvl <= max_lvl; lvl++) {
    adj_update_coarse(t, lvl, 0, m);
    adj_gsrb_boxes(t, lvl, t->n_down, m);
  }
}

int o_fas_vcycle_adj(otree *t, int max_lvl, const int *adj, double *vbar) {
  /* mg$D_fas_vcycle(tree, mg, .false., max_lvl) (m_aX_mg.f90:208-255)
   * transposed; adj and vbar as for o_fas_fmg_adj. Returns 0. */
  o_adjmg m = {adj[0], adj[1], adj[2], adj[3], vbar};
  for (int nb = 0; nb < t->nnb; nb++) vbar[nb] = 0;
  adj_vcycle(t, max_lvl, &m);
  return 0;
}

int o_fas_fmg_adj(otree *t, const int *adj, double *vbar) {
  /* mg$D_fas_fmg(tree, mg, .false., .false.) (m_aX_mg.f90:164-203)
   * transposed, in the reverse of the reference's schedule. adj = lp, lt,
   * lr, lw; vbar[nb-1] receives the adjoint of the value V of face nb.
   * The operators are those of the forward solve per box: the Laplacian
   * (with the Helmholtz term of o_mg_set_helmholtz), or the permittivity
   * forms for tagged boxes. Returns 0. */
  o_adjmg m = {adj[0], adj[1], adj[2], adj[3], vbar};
  for (int nb = 0; nb < t->nnb; nb++) vbar[nb] = 0;
  int min_lvl = t->min_lvl, max_lvl = t->max_lvl;
  for (int lvl = max_lvl; lvl >= min_lvl; lvl--) {
    adj_vcycle(t, lvl, &m);
    if (lvl > min_lvl) {
      adj_gc_level(t, lvl, &m);
      adj_correct_children(t, lvl - 1, &m);
    }
    ivec *ids = &t->ids[LV(t, lvl)];
    for (int q = ids->n - 1; q >= 0; q--) whole_adj(t, ids->v[q], 0, &m);  /* tmp = phi */
  }
  for (int lvl = min_lvl + 1; lvl <= max_lvl; lvl++) adj_update_coarse(t, lvl, 1, &m);
  return 0;
}

void o_store_vars(otree *t, int n_cc, const int *ivs, int n_fc, const int *fvs, double *buf,
                  int load) {
  /* whole cell and face variables of box slots 0..max_id to (load 0) or
   * from (load 1) buf, in the order given */
  size_t o = 0;
  for (int k = 0; k < n_cc; k++)
    for (int id = 0; id <= t->max_id; id++) {
      double *c = CC(t, id, ivs[k]);
      if (load) memcpy(c, buf + o, sizeof(double) * t->box_len);
      else memcpy(buf + o, c, sizeof(double) * t->box_len);
      o += t->box_len;
    }
  for (int k = 0; k < n_fc; k++)
    for (int id = 0; id <= t->max_id; id++)
      for (int d = 0; d < t->dim; d++) {
        double *f = FC(t, id, fvs[k], d);
        if (load) memcpy(f, buf + o, sizeof(double) * t->face_len);
        else memcpy(buf + o, f, sizeof(double) * t->face_len);
        o += t->face_len;
      }
}

size_t o_store_size(const otree *t, int n_cc, int n_fc) {
  return (size_t)(t->max_id + 1) * ((size_t)n_cc * t->box_len + (size_t)n_fc * t->dim * t->face_len);
}
