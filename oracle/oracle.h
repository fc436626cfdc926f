/* CPU oracle of the clustering hot path (test infrastructure only; see oracle.c). */
#ifndef CUHC_ORACLE_H
#define CUHC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

float oracle_distance(const float *a, const float *b, int d, int metric, float p);
int oracle_pairwise(const float *A, int nA, const float *B, int nB, int d, int metric, float p,
                    float *out);
int oracle_mst(const float *X, int n, int d, int metric, float p, const float *core, float alpha,
               int *eu, int *ev, float *ew);
int oracle_tree_from_mst(const int *eu, const int *ev, const float *ew, int n, double *tree);
int oracle_single_linkage(const float *X, int n, int d, int metric, float p, double *tree);
int oracle_core_distances(const float *X, int n, int d, int metric, float p, int k, float *core);
int oracle_hdbscan_tree(const float *X, int n, int d, int metric, float p, int k, float alpha,
                        double *tree, float *core);
int oracle_dense_on_matrix(float *D, int n, int linkage, double *tree);
int oracle_dense_linkage(const float *X, int n, int d, int metric, float p, int linkage,
                         double *tree);
int oracle_mlp_forward(const float *p, const float *X, int m, double *y);
int oracle_mlp_gradient(const float *p, const float *X, const float *t, int m, double *grad,
                        double *loss);

#ifdef __cplusplus
}
#endif

#endif
