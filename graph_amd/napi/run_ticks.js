// Runs ticks through the Node-API addon on inputs written as raw little-endian
// arrays in a directory (xy.f32, links.f32, rin.f32, rout.f32, jitter.f32,
// seeds.f32, params.json) and writes out.f32 (positions) and alpha.f64.
'use strict'
const fs = require('fs')
const path = require('path')
const addon = require(path.join(__dirname, 'cosmosgl.node'))

function f32 (dir, name) {
  const b = fs.readFileSync(path.join(dir, name))
  return new Float32Array(b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength))
}

const dir = process.argv[2]
const params = JSON.parse(fs.readFileSync(path.join(dir, 'params.json'), 'utf8'))
const h = addon.create(0, 0)
const xy = f32(dir, 'xy.f32')
addon.setPositions(h, xy)
addon.setJitter(h, f32(dir, 'jitter.f32'))
addon.setLinks(h, f32(dir, 'links.f32'), null, f32(dir, 'rin.f32'), f32(dir, 'rout.f32'))
addon.setAlpha(h, 1.0)
const seeds = f32(dir, 'seeds.f32')
const trace = addon.step(h, params, seeds.length, seeds)
const out = new Float32Array(xy.length)
addon.getPositions(h, out)
fs.writeFileSync(path.join(dir, 'out.f32'), Buffer.from(out.buffer))
fs.writeFileSync(path.join(dir, 'alpha.f64'), Buffer.from(trace.buffer))
addon.destroy(h)
console.log('ticks', seeds.length, 'alpha', trace[trace.length - 1])
