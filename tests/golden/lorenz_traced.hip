// Traced right-hand side (straight-line, operation order as recorded).
template <class T>
__device__ inline void lorenz_traced(T* du, const T* u, const T* p, const T* ps, double t) {
  const T v0 = u[1] - u[0];
  const T v1 = p[0] * v0;
  const T v2 = p[1] - u[2];
  const T v3 = u[0] * v2;
  const T v4 = v3 - u[1];
  const T v5 = u[0] * u[1];
  const T v6 = p[2] * u[2];
  const T v7 = v5 - v6;
  du[0] = v1;
  du[1] = v4;
  du[2] = v7;
}
