// Evaluates force passes through the Node-API addon on inputs written as raw
// little-endian arrays in a directory (xy.f32, links.f32, rin.f32, rout.f32,
// jitter.f32, sizes.f32 (optional), params.json with {params, alpha, seed,
// forces: [codes]}) and writes v<code>.f32 per requested pass.
'use strict'
const fs = require('fs')
const path = require('path')
const addon = require(path.join(__dirname, 'cosmosgl.node'))

function f32 (dir, name) {
  const file = path.join(dir, name)
  if (!fs.existsSync(file)) return null
  const b = fs.readFileSync(file)
  return new Float32Array(b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength))
}

const dir = process.argv[2]
const job = JSON.parse(fs.readFileSync(path.join(dir, 'params.json'), 'utf8'))
const h = addon.create(0, 0)
const xy = f32(dir, 'xy.f32')
addon.setPositions(h, xy)
addon.setJitter(h, f32(dir, 'jitter.f32'))
addon.setLinks(h, f32(dir, 'links.f32'), null, f32(dir, 'rin.f32'), f32(dir, 'rout.f32'))
const sizes = f32(dir, 'sizes.f32')
if (sizes) addon.setPointSizes(h, sizes)
for (const force of job.forces) {
  const v = addon.evalForce(h, job.params, force, job.alpha, job.seed, new Float32Array(xy.length))
  fs.writeFileSync(path.join(dir, `v${force}.f32`), Buffer.from(v.buffer))
}
addon.destroy(h)
console.log('passes', job.forces.length)
