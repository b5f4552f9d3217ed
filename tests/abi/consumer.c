/* C99 consumer of include/aurora_amd.h: the header compiles as C (-std=c99
 * -pedantic -Werror), its records have the reference's sizes and offsets
 * (src/util/shader.h:23-37, src/util/light.h:24-35, src/util/transform.h:22-25,
 * src/util/camera.h:30-35, src/util/emitter.h:13-18), and every entry point
 * links against libaurora_amd.so.  Prints "c-consumer ok <entry points>". */
#include <stddef.h>
#include <stdio.h>

#include "aurora_amd.h"

_Static_assert(sizeof(aur_shader) == 80, "Shader is 80 B");
_Static_assert(offsetof(aur_shader, diffuse_color) == 4, "Shader.diffuseColor");
_Static_assert(offsetof(aur_shader, emission_color) == 16, "Shader.emissionColor");
_Static_assert(offsetof(aur_shader, specular_color) == 28, "Shader.specularColor");
_Static_assert(offsetof(aur_shader, reflection_color) == 40, "Shader.reflectionColor");
_Static_assert(offsetof(aur_shader, diffuse) == 52, "Shader.diffuse");
_Static_assert(offsetof(aur_shader, emission) == 56, "Shader.emission");
_Static_assert(offsetof(aur_shader, exponent) == 60, "Shader.exponent");
_Static_assert(offsetof(aur_shader, reflectivity) == 64, "Shader.reflectivity");
_Static_assert(offsetof(aur_shader, translucence) == 68, "Shader.translucence");
_Static_assert(offsetof(aur_shader, refractive_index) == 72, "Shader.refractiveIndex");
_Static_assert(offsetof(aur_shader, texture) == 76, "Shader.texture");

_Static_assert(sizeof(aur_light) == 80, "Light is 80 B");
_Static_assert(offsetof(aur_light, intensity) == 4, "Light.intensity");
_Static_assert(offsetof(aur_light, ambient) == 8, "Light.ambient");
_Static_assert(offsetof(aur_light, area) == 12, "Light.area");
_Static_assert(offsetof(aur_light, samples) == 16, "Light.samples");
_Static_assert(offsetof(aur_light, position) == 20, "Light.position");
_Static_assert(offsetof(aur_light, color) == 32, "Light.color");
_Static_assert(offsetof(aur_light, direction) == 44, "Light.direction");
_Static_assert(offsetof(aur_light, e1) == 56, "Light.e1");
_Static_assert(offsetof(aur_light, e2) == 68, "Light.e2");

_Static_assert(sizeof(aur_transform) == 136, "Transform is 136 B");
_Static_assert(offsetof(aur_transform, size) == 4, "Transform.size");
_Static_assert(offsetof(aur_transform, world) == 8, "Transform.worldMatrix");
_Static_assert(offsetof(aur_transform, normal) == 72, "Transform.normalMatrix");

_Static_assert(offsetof(aur_camera, forward) == 12, "Camera.forward");
_Static_assert(offsetof(aur_camera, right) == 24, "Camera.right");
_Static_assert(offsetof(aur_camera, up) == 36, "Camera.up");
_Static_assert(offsetof(aur_camera, tanfov) == 48, "Camera.tanfov");
_Static_assert(offsetof(aur_camera, aspect) == 56, "Camera.aspect");

_Static_assert(sizeof(aur_emitter) == 28, "Emitter is 28 B");
_Static_assert(offsetof(aur_emitter, pdf) == 4, "Emitter.pdf");
_Static_assert(offsetof(aur_emitter, cdf) == 8, "Emitter.cdf");
_Static_assert(offsetof(aur_emitter, power) == 12, "Emitter.power");
_Static_assert(offsetof(aur_emitter, triangle) == 24, "Emitter.triangleID");

/* every entry point, so that the link fails if one is missing */
typedef void (*entry_fn)(void);
static const entry_fn entry_points[] = {
#define AUR_ENTRY(name) (entry_fn)&name,
#include "entry_points.inc"
#undef AUR_ENTRY
};

int main(void) {
  aur_context* ctx = NULL;
  const aur_status st = aur_create(0, &ctx);
  printf("c-consumer ok %u %d %u\n", (unsigned)(sizeof(entry_points) / sizeof(entry_points[0])), (int)st,
         aur_abi_version());
  if (ctx) aur_destroy(ctx);
  return 0;
}
