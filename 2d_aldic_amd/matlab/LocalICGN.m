function [U,F,HtempPar,LocalTime,ConvItPerEle,LocalICGNBadPtNum] = LocalICGN(U0,coordinatesFEM,...
    Df,ImgRef,ImgDef,DICpara,ICGNmethod,tol)
% Library-backed replacement of LocalICGN.m:1-161 (called main_ALDIC.m:178): 6-DOF IC-GN of
% every node (funICGN.m) on the GPU, the bad-point rule (status < 0 or > 99, then above
% mean + 0.15 std of the good ones) and the natural-neighbour refill of U and F
% (scatteredInterpolant 'natural') on the GPU.  Gauss-Newton only (the reference's LM
% branch calls a method plain image arrays lack).
if nargin < 8, tol = 1e-2; end
if nargin >= 7 && ~strcmp(ICGNmethod, 'GaussNewton')
    error('aldic:icgn', 'only GaussNewton is supported');
end
aldic_mex('set_normalized', double(ImgRef), double(ImgDef));
aldic_mex('set_gradients', Df.DfDx, Df.DfDy);
t0 = tic;
[U, F, H21, it] = aldic_mex('icgn6', U0(:), coordinatesFEM, DICpara.winsize, tol);
LocalTime = toc(t0);
HtempPar = H21';
ConvItPerEle = double(it);
[U, F, LocalICGNBadPtNum] = local_badpoints(U, F, ConvItPerEle, coordinatesFEM, 0);
end

function [U, F, nbad] = local_badpoints(U, F, it, coords, floor20)
% status < 0 or > 99 (102 = masked, not counted), then status above mean + 0.15 std of
% the good ones (at least 20 with floor20); bad U and F refilled by natural neighbours
bad = it < 0 | it > 99;
nbad = nnz(bad) - nnz(it == 102);
good = ~bad;
if any(good)
    thr = mean(it(good)) + 0.15 * std(it(good));
    if floor20, thr = max(thr, 20); end
    bad = bad | it > thr;
end
U([2*find(bad)-1; 2*find(bad)]) = NaN;
if ~isempty(F)
    idx = find(bad);
    F([4*idx-3; 4*idx-2; 4*idx-1; 4*idx]) = NaN;
end
[U, F] = aldic_mex('natural_fill', coords, bad, U, F);
end
