#pragma once
#include <stddef.h>

namespace zk {

struct Mlp;
int mlp_create(int nlayers, const int *dims, unsigned seed, Mlp **out);
void mlp_destroy(Mlp *n);
size_t mlp_nparams(const Mlp *n);
int mlp_layers(const Mlp *n);
int mlp_dim(const Mlp *n, int i);
int cost_model_set(Mlp *m);
void cost_model_forget(Mlp *m);  // uninstalls m if it is the installed cost model
int mlp_get_params(Mlp *n, float *host);
int mlp_set_params(Mlp *n, const float *host);
int mlp_forward(Mlp *n, int B, const float *X_host, float *Y_host);
int mlp_gradients(Mlp *n, int B, const float *X_host, const float *y_host, float *g_host, float *loss);
int mlp_train(Mlp *n, int B, const float *X_host, const float *y_host, int epochs, float lr, float *final_loss);

}  // namespace zk
