function [U,HPar,ALSub1Time,ConvItPerEle,LocalICGNBadPtNum] = Subpb1Quadtree(UOld,FOld,udual,vdual,coordinatesFEM,...
    Df,ImgRef,ImgDef,mu,beta,HPar,ALSolveStep,DICpara,ICGNmethod,tol)
% Library-backed replacement of Subpb1Quadtree.m:1-199: 2-DOF IC-GN on the GPU with the
% window list mean(winsize) at every node (:50), bad = status < 0 or > 99 or above
% max(mean + 0.15 std, 20) of the good ones (:170-181), bad U refilled by natural
% neighbours on the GPU (:188-199).
if nargin < 15, tol = 1e-2; end
S = size(coordinatesFEM, 1);
if iscell(HPar), H = zeros(21, S); for k = 1:21, H(k, :) = HPar{k}(:)'; end
else, H = HPar'; end
if ALSolveStep <= 1, H = zeros(21, S); end
aldic_mex('set_normalized', double(ImgRef), double(ImgDef));
aldic_mex('set_gradients', Df.DfDx, Df.DfDy);
t0 = tic;
[U, it] = aldic_mex('icgn2', coordinatesFEM, round(mean(DICpara.winsize)), H, mu, vdual(:), ...
    UOld(:), FOld(:), tol);
ALSub1Time = toc(t0);
ConvItPerEle = double(it);
bad = ConvItPerEle < 0 | ConvItPerEle > 99;
LocalICGNBadPtNum = nnz(bad) - nnz(ConvItPerEle == 102);
good = ~bad;
if any(good)
    bad = bad | ConvItPerEle > max(mean(ConvItPerEle(good)) + 0.15 * std(ConvItPerEle(good)), 20);
end
U([2*find(bad)-1; 2*find(bad)]) = NaN;
[U, ~] = aldic_mex('natural_fill', coordinatesFEM, bad, U, []);
end
