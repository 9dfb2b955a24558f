__kernel void probe(__global const float* a, __global const float* b, __global float* out, int n) {
  int i = get_global_id(0);
  if (i >= n) return;
  float3 x = vload3(i, a);
  float3 y = vload3(i, b);
  float d = dot(x, y);
  float3 c = cross(x, y);
  float l = length(x);
  float3 nz = normalize(x);
  float q = x.x / y.x;
  out[9*i+0] = d; out[9*i+1] = c.x; out[9*i+2] = c.y; out[9*i+3] = c.z;
  out[9*i+4] = l; out[9*i+5] = nz.x; out[9*i+6] = nz.y; out[9*i+7] = nz.z; out[9*i+8] = q;
}
